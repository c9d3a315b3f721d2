// QEq: per-edge Coulomb-kernel values and a fused two-right-hand-side CG.
//
// The reference assembles the dense KKT system [[H, 1], [1^T, 0]] and calls
// LAPACK (src/xreac/energy.py:89-103), with
//   H_ij = 14.4 * sum_S shield(r_ijS) + delta_ij * eta_i.
// Here H stays implicit in the canonical edge list: every directed edge e of
// row i carries h_e = 14.4 * shield(r_e), so (H q)_i = eta_i q_i + sum_e h_e q_j(e)
// (self-image edges e = (i, i, S != 0) land on the diagonal exactly as
// pair_sum does).  The constrained minimum is q = s + mu t with H s = -chi,
// H t = -1, mu = (Q - sum s) / sum t; s and t share every sweep over the
// edges (one 16-byte gather of (p_s, p_t) per edge).
//
// One CG iteration is two kernels.  Each consumer block re-reduces the
// producer's per-block partials in a fixed order, so no separate reduction
// launches are needed and the result is deterministic.
#include "launch.h"
#include "qeq.h"
#include "profile.h"

namespace {

constexpr int ROWS_PER_BLOCK = 4;  // one wave per row, 256 threads

// ---------------------------------------------------------------- edge pass
__global__ void __launch_bounds__(256) k_edge_prep(int n, const double* __restrict__ x, const int* __restrict__ type,
                                                   CellDev cell, FFDev ff, const int* __restrict__ row_ptr,
                                                   const int* __restrict__ nbr_j, const int* __restrict__ nbr_s,
                                                   double bond_cut, double* __restrict__ hval,
                                                   unsigned char* __restrict__ bflag, int* __restrict__ bond_count) {
  const int i = blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int ti = type[i];
  const double rc = ff.gen[12];
  const double cut = 0.01 * ff.gen[29];
  int cnt = 0;
  for (int e = row_ptr[i] + lane; e < row_ptr[i + 1]; e += 64) {
    const int j = nbr_j[e];
    const int tj = type[j];
    double v[3];
    const double r = xr_edge(x, cell, i, j, nbr_s[e], v);
    double s, ds;
    xr_shield(r, rc, ff.pair[(ti * ff.nt + tj) * PP_COUNT + PP_GAMMA], s, ds);
    hval[e] = 14.4 * s;
    unsigned char active = 0;
    if (r <= bond_cut) {
      BondRaw b;
      xr_bond_raw(ff, ti, tj, r, b);
      active = ((b.sig + b.pi) + b.pp) >= cut;
    }
    bflag[e] = active;
    cnt += active;
  }
  cnt = xr_wave_sum_int(cnt);
  if (lane == 0) bond_count[i] = cnt;
}

// Compact active edges per row, preserving canonical order.
__global__ void __launch_bounds__(256) k_bond_fill(int n, const int* __restrict__ row_ptr,
                                                   const unsigned char* __restrict__ bflag,
                                                   const int* __restrict__ bond_ptr, int cap,
                                                   int* __restrict__ bond_edge, int* __restrict__ bond_i) {
  const int i = blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  int base = bond_ptr[i];
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int e0 = row_ptr[i]; e0 < row_ptr[i + 1]; e0 += 64) {
    const int e = e0 + lane;
    const bool on = e < row_ptr[i + 1] && bflag[e];
    const unsigned long long m = __ballot(on);
    if (on) {
      const int b = base + __popcll(m & below);
      if (b < cap) {  // pointers are clamped to the buffer; overflow is detected by the host
        bond_edge[b] = e;
        bond_i[b] = i;
      }
    }
    base += __popcll(m);
  }
}

// --------------------------------------------------------------------- CG
// Kernel boundaries order all cross-block communication: each kernel writes
// per-block partial dot products, and every block of the next kernel reduces
// them again in the same fixed order (redundant but cheap, and bitwise
// reproducible).  Device-scope fences inside a kernel are avoided on purpose:
// on a multi-XCD part they write back the whole L2.
struct CgScalars {
  int done, iters;
};

constexpr int UPD_BLOCK = 256;
constexpr int DIR_BLOCK = 1024;     // 16 waves, one row per wave at a time
constexpr int DIR_MAX_BLOCKS = 256;  // 4096 waves in flight; few partial sums to reduce

// Fixed-order reduction of NV partial arrays ([k * nb + b]) by the whole block.
template <int NV>
__device__ inline void reduce_partials(const double* __restrict__ part, int nb, double (&out)[NV]) {
  __shared__ double red[NV][16];
  __shared__ double res[NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
#pragma unroll 4
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[k * nb + b];
    s = xr_wave_sum(s);
    if (lane == 0) red[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[threadIdx.x][w];
    res[threadIdx.x] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = res[k];
  __syncthreads();
}

__global__ void __launch_bounds__(UPD_BLOCK) k_cg_init(int n, const int* __restrict__ type, FFDev ff,
                                                       double2* __restrict__ X, double2* __restrict__ R,
                                                       double4* __restrict__ ZP, double* __restrict__ inv_diag,
                                                       double* __restrict__ eta_out, double* __restrict__ part_b,
                                                       double* __restrict__ part_bb, CgScalars* __restrict__ sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double vals[4] = {0, 0, 0, 0};
  double bb[2] = {0, 0};
  if (i < n) {
    const double* a = ff.atom + type[i] * AP_COUNT;
    const double eta = a[AP_ETA];
    const double d = 1.0 / eta;
    const double2 r = make_double2(-a[AP_CHI], -1.0);
    X[i] = make_double2(0.0, 0.0);
    R[i] = r;
    const double2 z = make_double2(r.x * d, r.y * d);
    ZP[i] = make_double4(z.x, z.y, 0.0, 0.0);  // (z_s, z_t, p_s, p_t), p = 0 before iteration 0
    inv_diag[i] = d;
    eta_out[i] = eta;
    vals[0] = r.x * z.x;
    vals[1] = r.y * z.y;
    vals[2] = r.x * r.x;
    vals[3] = r.y * r.y;
    bb[0] = vals[2];
    bb[1] = vals[3];
  }
  // slot 1 holds the residual state before iteration 0
  xr_block_partials<4>(vals, part_b + 4 * gridDim.x, gridDim.x);
  xr_block_partials<2>(bb, part_bb, gridDim.x);
  if (i == 0) {
    sc->done = 0;
    sc->iters = -1;
  }
}

// Iteration k, first half: beta, p = z + beta p_old and A p (grid-stride rows,
// one wave per row), partial p.Ap per block.  zp_in holds (z, p_old) of every
// atom; the new p goes to the p half of zp_out.  Up to 512 edges of a row are
// loaded before any gather so a row costs one memory latency.
__global__ void __launch_bounds__(DIR_BLOCK) k_cg_direction(int n, int k, const int* __restrict__ row_ptr,
                                                      const int* __restrict__ nbr_j, const double* __restrict__ hval,
                                                      const double* __restrict__ eta, const double4* __restrict__ zp_in,
                                                      double4* __restrict__ zp_out, double2* __restrict__ AP,
                                                      const double* __restrict__ part_b,
                                                      const double* __restrict__ part_bb, int nb,
                                                      double* __restrict__ part_a, double tol2,
                                                      CgScalars* __restrict__ sc) {
  if (sc->done) return;
  double cur[4], bb[2];
  reduce_partials<4>(part_b + ((k - 1) & 1) * 4 * nb, nb, cur);
  reduce_partials<2>(part_bb, nb, bb);
  const bool conv_s = cur[2] <= tol2 * bb[0];
  const bool conv_t = cur[3] <= tol2 * bb[1];
  if (conv_s && conv_t) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->done = 1;
      sc->iters = k;
    }
    return;
  }
  double beta_s = 0.0, beta_t = 0.0;
  if (k > 0) {
    double prev[4];
    reduce_partials<4>(part_b + (k & 1) * 4 * nb, nb, prev);
    if (!conv_s && prev[0] != 0.0) beta_s = cur[0] / prev[0];
    if (!conv_t && prev[1] != 0.0) beta_t = cur[1] / prev[1];
  }
  const int lane = threadIdx.x & 63;
  const int waves = gridDim.x * (blockDim.x >> 6);
  double pap_s = 0.0, pap_t = 0.0;
  for (int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += waves) {
    double acc_s = 0.0, acc_t = 0.0;
    const int e0 = row_ptr[i], e1 = row_ptr[i + 1];
    for (int base = e0; base < e1; base += 512) {
      int jj[8];
      double hh[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = base + u * 64 + lane;
        const bool ok = e < e1;
        jj[u] = ok ? nbr_j[e] : i;
        hh[u] = ok ? hval[e] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double4 v = zp_in[jj[u]];
        acc_s += hh[u] * (v.x + beta_s * v.z);
        acc_t += hh[u] * (v.y + beta_t * v.w);
      }
    }
    acc_s = xr_wave_sum(acc_s);
    acc_t = xr_wave_sum(acc_t);
    if (lane == 0) {
      const double4 v = zp_in[i];
      const double2 p = make_double2(v.x + beta_s * v.z, v.y + beta_t * v.w);
      const double d = eta[i];
      const double2 ap = make_double2(d * p.x + acc_s, d * p.y + acc_t);
      reinterpret_cast<double2*>(zp_out + i)[1] = p;
      AP[i] = ap;
      pap_s += p.x * ap.x;
      pap_t += p.y * ap.y;
    }
  }
  double vals[2] = {pap_s, pap_t};
  xr_block_partials<2>(vals, part_a, gridDim.x);
}

// Iteration k, second half: alpha, x += alpha p, r -= alpha A p, z = r / diag
// (into the z half of zp), partial r.z and r.r per block.
__global__ void __launch_bounds__(UPD_BLOCK) k_cg_update(int n, int k, const double* __restrict__ inv_diag,
                                                         double2* __restrict__ X, double2* __restrict__ R,
                                                         double4* __restrict__ zp, const double2* __restrict__ AP,
                                                         const double* __restrict__ part_a, int na,
                                                         double* __restrict__ part_b,
                                                         const double* __restrict__ part_bb, double tol2,
                                                         const CgScalars* __restrict__ sc) {
  if (sc->done) return;
  const int nb = gridDim.x;
  double cur[4], pap[2], bb[2];
  reduce_partials<4>(part_b + ((k - 1) & 1) * 4 * nb, nb, cur);
  reduce_partials<2>(part_a, na, pap);
  reduce_partials<2>(part_bb, nb, bb);
  const bool conv_s = cur[2] <= tol2 * bb[0];
  const bool conv_t = cur[3] <= tol2 * bb[1];
  const double alpha_s = (!conv_s && pap[0] != 0.0) ? cur[0] / pap[0] : 0.0;
  const double alpha_t = (!conv_t && pap[1] != 0.0) ? cur[1] / pap[1] : 0.0;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double vals[4] = {0, 0, 0, 0};
  if (i < n) {
    const double2 p = reinterpret_cast<const double2*>(zp + i)[1];
    const double2 ap = AP[i];
    double2 xv = X[i], r = R[i];
    xv.x += alpha_s * p.x;
    xv.y += alpha_t * p.y;
    r.x -= alpha_s * ap.x;
    r.y -= alpha_t * ap.y;
    const double d = inv_diag[i];
    const double2 z = make_double2(r.x * d, r.y * d);
    X[i] = xv;
    R[i] = r;
    reinterpret_cast<double2*>(zp + i)[0] = z;
    vals[0] = r.x * z.x;
    vals[1] = r.y * z.y;
    vals[2] = r.x * r.x;
    vals[3] = r.y * r.y;
  }
  xr_block_partials<4>(vals, part_b + (k & 1) * 4 * nb, nb);
}

// q = s + mu t with mu = (Q - sum s) / sum t.
__global__ void __launch_bounds__(1024) k_qeq_charges(int n, const double2* __restrict__ X, double total_charge,
                                                      double* __restrict__ q, double* __restrict__ mu_out) {
  __shared__ double red[2][16];
  double ss = 0.0, st = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    ss += X[i].x;
    st += X[i].y;
  }
  ss = xr_wave_sum(ss);
  st = xr_wave_sum(st);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][w] = ss;
    red[1][w] = st;
  }
  __syncthreads();
  __shared__ double mu_s;
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
      a += red[0][k];
      b += red[1][k];
    }
    mu_s = (total_charge - a) / b;
    mu_out[0] = mu_s;
  }
  __syncthreads();
  const double mu = mu_s;
  for (int i = threadIdx.x; i < n; i += blockDim.x) q[i] = X[i].x + mu * X[i].y;
}

}  // namespace

hipError_t xr_edge_prep(int n, const double* x, const int* type, CellDev cell, FFDev ff, const int* row_ptr,
                        const int* nbr_j, const int* nbr_s, double bond_cut, double* hval, unsigned char* bflag,
                        int* bond_count, hipStream_t s) {
  const int blocks = (n + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;
  XR_LAUNCH(XT_EDGE, k_edge_prep, dim3(blocks), dim3(256), 0, s, n, x, type, cell, ff, row_ptr, nbr_j, nbr_s,
            bond_cut, hval, bflag, bond_count);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_bond_fill(int n, const int* row_ptr, const unsigned char* bflag, const int* bond_ptr, int cap,
                        int* bond_edge, int* bond_i, hipStream_t s) {
  const int blocks = (n + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;
  hipLaunchKernelGGL(k_bond_fill, dim3(blocks), dim3(256), 0, s, n, row_ptr, bflag, bond_ptr, cap, bond_edge,
                     bond_i);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

size_t xr_cg_partials(int n) {
  const int nb = (n + UPD_BLOCK - 1) / UPD_BLOCK;
  return (size_t)(2 * 4 * nb + 2 * nb + 2 * DIR_MAX_BLOCKS + n) + 64;
}

size_t xr_cg_state_bytes() { return sizeof(CgScalars); }

hipError_t xr_cg_solve(const CgArgs& a, hipStream_t s, int* iterations) {
  const int n = a.n;
  const int nb = (n + UPD_BLOCK - 1) / UPD_BLOCK;
  int na = (n + DIR_BLOCK / 64 - 1) / (DIR_BLOCK / 64);
  if (na > DIR_MAX_BLOCKS) na = DIR_MAX_BLOCKS;
  double* part_b = a.partials;
  double* part_bb = part_b + 2 * 4 * nb;
  double* part_a = part_bb + 2 * nb;
  double* eta = part_a + 2 * DIR_MAX_BLOCKS;
  CgScalars* sc = static_cast<CgScalars*>(a.state);
  const double tol2 = a.tol * a.tol;
  hipLaunchKernelGGL(k_cg_init, dim3(nb), dim3(UPD_BLOCK), 0, s, n, a.type, a.ff, a.X, a.R, a.ZP0, a.inv_diag,
                     eta, part_b, part_bb, sc);
  XR_LAUNCH_CHECK();
  int k = 0;
  int host_state[3] = {0, -1, 0};
  bool first = true, stop = false;
  while (!stop) {
    const int chunk = first ? a.first_check : a.check_every;
    first = false;
    for (int c = 0; c < chunk; ++c) {
      double4* zp_in = (k & 1) ? a.ZP1 : a.ZP0;
      double4* zp_out = (k & 1) ? a.ZP0 : a.ZP1;
      XR_LAUNCH(XT_CG_DIRECTION, k_cg_direction, dim3(na), dim3(DIR_BLOCK), 0, s, n, k, a.row_ptr, a.nbr_j, a.hval, eta,
                zp_in, zp_out, a.AP, part_b, part_bb, nb, part_a, tol2, sc);
      if (k == a.max_iter) {  // this launch only tests the final residual
        stop = true;
        break;
      }
      XR_LAUNCH(XT_CG_UPDATE, k_cg_update, dim3(nb), dim3(UPD_BLOCK), 0, s, n, k, a.inv_diag, a.X, a.R, zp_out,
                a.AP, part_a, na, part_b, part_bb, tol2, sc);
      ++k;
    }
    XR_LAUNCH_CHECK();
    hipError_t e = hipMemcpyAsync(host_state, sc, 2 * sizeof(int), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    if (a.abort_flag) {
      e = hipMemcpyAsync(host_state + 2, a.abort_flag, sizeof(int), hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) return e;
    }
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    if (host_state[2]) {  // an input error was flagged upstream; stop iterating
      *iterations = -2;
      return hipSuccess;
    }
    if (host_state[0]) break;
  }
  *iterations = host_state[0] ? host_state[1] : -1;
  hipLaunchKernelGGL(k_qeq_charges, dim3(1), dim3(1024), 0, s, n, a.X, a.total_charge, a.q, a.mu);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}
