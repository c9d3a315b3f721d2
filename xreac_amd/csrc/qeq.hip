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
                                                   const int* __restrict__ bond_ptr, int* __restrict__ bond_edge,
                                                   int* __restrict__ bond_i) {
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
      bond_edge[b] = e;
      bond_i[b] = i;
    }
    base += __popcll(m);
  }
}

// --------------------------------------------------------------------- CG
// Scalars shared by the CG kernels (device memory).  Each kernel's per-block
// partial dot products are reduced once, in a fixed order, by whichever block
// finishes last (integer atomic ticket), which then writes alpha or beta.
struct CgScalars {
  double rz[2], rr[2], bb[2], alpha[2], beta[2];
  int done, iters;
  unsigned ticket;
  int pad;
};

__device__ inline double load_agent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// True in exactly one block: the last to arrive.  Partials stored by any
// block before the call are visible to it.
__device__ inline bool last_block(unsigned* ticket) {
  __shared__ bool last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  __syncthreads();
  if (last) __threadfence();
  return last;
}

// Fixed-order sum of NV partial arrays ([k * nb + b]) by the first wave.
template <int NV>
__device__ inline void reduce_partials(const double* part, int nb, double (&out)[NV]) {
  __shared__ double res[NV];
  if (threadIdx.x < 64) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int b = threadIdx.x; b < nb; b += 64) s += load_agent(part + k * nb + b);
      s = xr_wave_sum(s);
      if (threadIdx.x == 0) res[k] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = res[k];
}

constexpr int UPD_BLOCK = 128;

__global__ void __launch_bounds__(UPD_BLOCK) k_cg_init(int n, const int* __restrict__ type, FFDev ff,
                                                       double2* __restrict__ X, double2* __restrict__ R,
                                                       double2* __restrict__ Z, double* __restrict__ inv_diag,
                                                       double* __restrict__ eta_out, double* __restrict__ part,
                                                       double tol2, CgScalars* __restrict__ sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double vals[4] = {0, 0, 0, 0};
  if (i < n) {
    const double* a = ff.atom + type[i] * AP_COUNT;
    const double eta = a[AP_ETA];
    const double d = 1.0 / eta;
    const double2 r = make_double2(-a[AP_CHI], -1.0);
    X[i] = make_double2(0.0, 0.0);
    R[i] = r;
    const double2 z = make_double2(r.x * d, r.y * d);
    Z[i] = z;
    inv_diag[i] = d;
    eta_out[i] = eta;
    vals[0] = r.x * z.x;
    vals[1] = r.y * z.y;
    vals[2] = r.x * r.x;
    vals[3] = r.y * r.y;
  }
  xr_block_partials<4>(vals, part, gridDim.x);
  if (last_block(&sc->ticket)) {
    double s[4];
    reduce_partials<4>(part, gridDim.x, s);
    if (threadIdx.x == 0) {
      for (int k = 0; k < 2; ++k) {
        sc->rz[k] = s[k];
        sc->rr[k] = s[2 + k];
        sc->bb[k] = s[2 + k];  // b = r at x = 0
        sc->beta[k] = 0.0;
        sc->alpha[k] = 0.0;
      }
      const bool conv = s[2] <= tol2 * s[2] && s[3] <= tol2 * s[3];
      sc->done = conv ? 1 : 0;
      sc->iters = conv ? 0 : -1;
      sc->ticket = 0;
    }
  }
}

// Iteration k, first half: p = z + beta p_old, A p, and alpha = (r.z)/(p.Ap).
__global__ void __launch_bounds__(256) k_cg_direction(int n, int k, const int* __restrict__ row_ptr,
                                                      const int* __restrict__ nbr_j, const double* __restrict__ hval,
                                                      const double* __restrict__ eta, const double2* __restrict__ Z,
                                                      const double2* __restrict__ Pold, double2* __restrict__ Pnew,
                                                      double2* __restrict__ AP, double* __restrict__ part,
                                                      double tol2, CgScalars* __restrict__ sc) {
  if (sc->done) return;
  const double beta_s = sc->beta[0], beta_t = sc->beta[1];
  const int i = blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  double vals[2] = {0.0, 0.0};
  if (i < n) {
    double as0 = 0.0, at0 = 0.0, as1 = 0.0, at1 = 0.0;
    const int e0 = row_ptr[i], e1 = row_ptr[i + 1];
    int e = e0 + lane;
    if (k > 0) {
      for (; e + 64 < e1; e += 128) {
        const int ja = nbr_j[e], jb = nbr_j[e + 64];
        const double ha = hval[e], hb = hval[e + 64];
        const double2 za = Z[ja], zb = Z[jb], pa = Pold[ja], pb = Pold[jb];
        as0 += ha * (za.x + beta_s * pa.x);
        at0 += ha * (za.y + beta_t * pa.y);
        as1 += hb * (zb.x + beta_s * pb.x);
        at1 += hb * (zb.y + beta_t * pb.y);
      }
      if (e < e1) {
        const int ja = nbr_j[e];
        const double ha = hval[e];
        const double2 za = Z[ja], pa = Pold[ja];
        as0 += ha * (za.x + beta_s * pa.x);
        at0 += ha * (za.y + beta_t * pa.y);
      }
    } else {
      for (; e + 64 < e1; e += 128) {
        const int ja = nbr_j[e], jb = nbr_j[e + 64];
        const double ha = hval[e], hb = hval[e + 64];
        const double2 za = Z[ja], zb = Z[jb];
        as0 += ha * za.x;
        at0 += ha * za.y;
        as1 += hb * zb.x;
        at1 += hb * zb.y;
      }
      if (e < e1) {
        const int ja = nbr_j[e];
        const double ha = hval[e];
        const double2 za = Z[ja];
        as0 += ha * za.x;
        at0 += ha * za.y;
      }
    }
    const double acc_s = xr_wave_sum(as0 + as1);
    const double acc_t = xr_wave_sum(at0 + at1);
    if (lane == 0) {
      const double2 z = Z[i];
      const double2 po = k > 0 ? Pold[i] : make_double2(0.0, 0.0);
      const double2 p = make_double2(z.x + beta_s * po.x, z.y + beta_t * po.y);
      const double d = eta[i];
      const double2 ap = make_double2(d * p.x + acc_s, d * p.y + acc_t);
      Pnew[i] = p;
      AP[i] = ap;
      vals[0] = p.x * ap.x;
      vals[1] = p.y * ap.y;
    }
  }
  xr_block_partials<2>(vals, part, gridDim.x);
  if (last_block(&sc->ticket)) {
    double pap[2];
    reduce_partials<2>(part, gridDim.x, pap);
    if (threadIdx.x == 0) {
      for (int c = 0; c < 2; ++c) {
        const bool conv = sc->rr[c] <= tol2 * sc->bb[c];
        sc->alpha[c] = (!conv && pap[c] != 0.0) ? sc->rz[c] / pap[c] : 0.0;
      }
      sc->ticket = 0;
    }
  }
}

// Iteration k, second half: x += alpha p, r -= alpha A p, z = r / diag,
// then beta, the residual test and the done flag.
__global__ void __launch_bounds__(UPD_BLOCK) k_cg_update(int n, int k, const double* __restrict__ inv_diag,
                                                         double2* __restrict__ X, double2* __restrict__ R,
                                                         double2* __restrict__ Z, const double2* __restrict__ P,
                                                         const double2* __restrict__ AP, double* __restrict__ part,
                                                         double tol2, CgScalars* __restrict__ sc) {
  if (sc->done) return;
  const double alpha_s = sc->alpha[0], alpha_t = sc->alpha[1];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double vals[4] = {0, 0, 0, 0};
  if (i < n) {
    const double2 p = P[i], ap = AP[i];
    double2 xv = X[i], r = R[i];
    xv.x += alpha_s * p.x;
    xv.y += alpha_t * p.y;
    r.x -= alpha_s * ap.x;
    r.y -= alpha_t * ap.y;
    const double d = inv_diag[i];
    const double2 z = make_double2(r.x * d, r.y * d);
    X[i] = xv;
    R[i] = r;
    Z[i] = z;
    vals[0] = r.x * z.x;
    vals[1] = r.y * z.y;
    vals[2] = r.x * r.x;
    vals[3] = r.y * r.y;
  }
  xr_block_partials<4>(vals, part, gridDim.x);
  if (last_block(&sc->ticket)) {
    double s[4];
    reduce_partials<4>(part, gridDim.x, s);
    if (threadIdx.x == 0) {
      bool all = true;
      for (int c = 0; c < 2; ++c) {
        const bool conv = s[2 + c] <= tol2 * sc->bb[c];
        all = all && conv;
        sc->beta[c] = (!conv && sc->rz[c] != 0.0) ? s[c] / sc->rz[c] : 0.0;
        sc->rz[c] = s[c];
        sc->rr[c] = s[2 + c];
      }
      if (all) {
        sc->done = 1;
        sc->iters = k + 1;
      }
      sc->ticket = 0;
    }
  }
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
  XR_TIMED(XT_EDGE, s,
           hipLaunchKernelGGL(k_edge_prep, dim3(blocks), dim3(256), 0, s, n, x, type, cell, ff, row_ptr, nbr_j,
                              nbr_s, bond_cut, hval, bflag, bond_count));
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_bond_fill(int n, const int* row_ptr, const unsigned char* bflag, const int* bond_ptr, int* bond_edge,
                        int* bond_i, hipStream_t s) {
  const int blocks = (n + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;
  hipLaunchKernelGGL(k_bond_fill, dim3(blocks), dim3(256), 0, s, n, row_ptr, bflag, bond_ptr, bond_edge, bond_i);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

size_t xr_cg_partials(int n) {
  const int nb = (n + UPD_BLOCK - 1) / UPD_BLOCK;
  const int na = (n + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;
  return (size_t)(4 * nb + 2 * na + 2 * n) + 64;
}

size_t xr_cg_state_bytes() { return sizeof(CgScalars); }

hipError_t xr_cg_solve(const CgArgs& a, hipStream_t s, int* iterations) {
  const int n = a.n;
  const int nb = (n + UPD_BLOCK - 1) / UPD_BLOCK;
  const int na = (n + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;
  double* part_b = a.partials;
  double* part_a = part_b + 4 * nb;
  double* eta = part_a + 2 * na;
  CgScalars* sc = static_cast<CgScalars*>(a.state);
  const double tol2 = a.tol * a.tol;
  hipError_t e = hipMemsetAsync(&sc->ticket, 0, sizeof(unsigned), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cg_init, dim3(nb), dim3(UPD_BLOCK), 0, s, n, a.type, a.ff, a.X, a.R, a.Z, a.inv_diag, eta,
                     part_b, tol2, sc);
  XR_LAUNCH_CHECK();
  int k = 0;
  int host_state[2] = {0, -1};
  while (k < a.max_iter) {
    const int chunk = k == 0 ? a.first_check : a.check_every;
    for (int c = 0; c < chunk && k < a.max_iter; ++c, ++k) {
      double2* pnew = (k & 1) ? a.P1 : a.P0;
      double2* pold = (k & 1) ? a.P0 : a.P1;
      XR_TIMED(XT_CG_DIRECTION, s,
               hipLaunchKernelGGL(k_cg_direction, dim3(na), dim3(256), 0, s, n, k, a.row_ptr, a.nbr_j, a.hval, eta,
                                  a.Z, pold, pnew, a.AP, part_a, tol2, sc));
      XR_TIMED(XT_CG_UPDATE, s,
               hipLaunchKernelGGL(k_cg_update, dim3(nb), dim3(UPD_BLOCK), 0, s, n, k, a.inv_diag, a.X, a.R, a.Z,
                                  pnew, a.AP, part_b, tol2, sc));
    }
    XR_LAUNCH_CHECK();
    e = hipMemcpyAsync(host_state, &sc->done, sizeof(host_state), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    if (host_state[0]) break;
  }
  *iterations = host_state[0] ? host_state[1] : -1;
  hipLaunchKernelGGL(k_qeq_charges, dim3(1), dim3(1024), 0, s, n, a.X, a.total_charge, a.q, a.mu);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}
