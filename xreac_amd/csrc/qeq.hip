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
//
// Batches: every system owns a fixed set of blocks in both CG kernels (sized
// from its atom count only), reduces only its own partials and stops on its
// own convergence test, so a system's charges are bitwise the same alone and
// in any batch.
//
// Mixed precision: the A p sweep reads FP32 copies of H and of the gathered
// (z, p) records while x, r, z, p and every dot product stay FP64.  When the
// recurrence residual has dropped by the inner factor, the system's next
// direction launch replaces it by the FP64 true residual r = b - A x (one
// FP64 sweep over its rows), the update launch after it reduces the new
// residual, and the recurrence restarts from there (iterative refinement); a
// system is
// converged only when a true residual meets the tolerance.  A warm start
// begins with the same refresh.  All of it runs on the device: the host only
// polls how many systems have settled.
#include <cstdlib>
#include <type_traits>

#include "launch.h"
#include "qeq.h"
#include "profile.h"

namespace {

constexpr int ROWS_PER_BLOCK = 4;  // one wave per row, 256 threads
constexpr double BLOCK_BOND = 0.3;  // raw bond order joining consecutive atoms into one CG block
constexpr int RUN_MAX = 16;         // atoms per bonded run (molecule) of the CG preconditioner
constexpr int BLOCK_MAX = 16;       // atoms per CG preconditioner block (whole runs packed together)

// ---------------------------------------------------------------- edge pass
__global__ void __launch_bounds__(256) k_edge_prep(int n, const double* __restrict__ x, const int* __restrict__ type,
                                                   Systems S, FFDev ff, const int* __restrict__ row_ptr,
                                                   const int* __restrict__ nbr_j, const int* __restrict__ nbr_s,
                                                   double bond_cut, double* __restrict__ hval,
                                                   float* __restrict__ hval32, unsigned char* __restrict__ bflag,
                                                   int* __restrict__ bond_count, unsigned char* __restrict__ join) {
  const int i = blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const CellDev& cell = S.sys[__builtin_amdgcn_readfirstlane(S.asys[i])].cell;
  const int ti = type[i];
  const double rc = ff.gen[12];
  const double cut = 0.01 * ff.gen[29];
  const double hcut = fmin(7.5, rc);
  const double hb_i = ff.atom[ti * AP_COUNT + AP_HBOND];
  int cnt = 0;
  unsigned jm = 0;  // bit d-1: strong same-image bond to atom i - d (d = 1..3), for the CG blocks
  const int zero = xr_pack_shift(0, 0, 0);
  for (int e = row_ptr[i] + lane; e < row_ptr[i + 1]; e += 64) {
    const int j = nbr_j[e];
    const int tj = type[j];
    double v[3];
    const double r = xr_edge(x, cell, i, j, nbr_s[e], v);
    double s, ds;
    xr_shield(r, rc, ff.pair[(ti * ff.nt + tj) * PP_COUNT + PP_GAMMA], s, ds);
    hval[e] = 14.4 * s;
    if (hval32) hval32[e] = (float)(14.4 * s);
    unsigned char active = 0;
    if (r <= bond_cut) {
      BondRaw b;
      xr_bond_raw(ff, ti, tj, r, b);
      const double raw = (b.sig + b.pi) + b.pp;
      active = raw >= cut;
      if (raw >= BLOCK_BOND && j < i && j >= i - 3 && nbr_s[e] == zero) jm |= 1u << (i - j - 1);
    }
    unsigned char flag = active ? EF_BOND : 0;
    if (r <= hcut) {
      const double hb_j = ff.atom[tj * AP_COUNT + AP_HBOND];
      if (hb_i == 1.0 && hb_j == 2.0) flag |= EF_HB_DONOR;
      if (hb_i == 2.0 && hb_j == 1.0) flag |= EF_HB_ACCEPTOR;
    }
    bflag[e] = flag;
    cnt += active;
  }
  cnt = xr_wave_sum_int(cnt);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) jm |= __shfl_down(jm, off, 64);
  if (lane == 0) {
    bond_count[i] = cnt;
    join[i] = (unsigned char)jm;
  }
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
    const bool on = e < row_ptr[i + 1] && (bflag[e] & EF_BOND);
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
// its system's partials again in the same fixed order (redundant but cheap,
// and bitwise reproducible).  Device-scope fences inside a kernel are avoided
// on purpose: on a multi-XCD part they write back the whole L2.
constexpr int UPD_BLOCK = 256;      // rows per update block
constexpr int DIR_BLOCK = 1024;     // 16 waves, one row per wave at a time
constexpr int DIR_MAX_BLOCKS = 256;  // per system: 4096 waves; few partial sums to reduce
// Rows per wave a direction block aims for: a block's prologue and partials
// cost the same whatever its rows, so small systems get few, full blocks.
constexpr int DIR_ROWS_PER_WAVE = 3;
constexpr int INNER_MAX = 64;        // mixed precision: iterations between FP64 restarts
constexpr double INNER_DROP2 = 1e-10;  // mixed precision: squared residual drop before a restart
constexpr int MAX_RESTARTS = 12;

// System owning global block b (first-block offsets ascend with the system index).
__device__ inline int block_system(const Systems& S, int b, bool dir) {
  int lo = 0, hi = S.ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int start = dir ? S.sys[mid].dir0 : S.sys[mid].upd0;
    if (start <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Sums of NV series of cnt block partials in a fixed order: thread-strided
// loads over the whole block (one latency), the wave tree in the waves that
// hold data, then those waves in order.  One barrier; the order depends only
// on cnt and the block size, so only on the system.
template <int NV, class Load>
__device__ inline void block_sums(int cnt, Load load, double (&out)[NV]) {
  __shared__ double red[NV][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int used = (cnt + 63) >> 6 < (int)(blockDim.x >> 6) ? (cnt + 63) >> 6 : (int)(blockDim.x >> 6);
  if (wave < used) {
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.0;
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) load(t, v);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const double w = xr_wave_sum(v[q]);
      if (lane == 0) red[q][wave] = w;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double x = 0.0;
    for (int w = 0; w < used; ++w) x += red[q][w];
    out[q] = x;
  }
}

// The direction kernel's prologue: current and previous residual dots (4 + 4)
// and |b|^2 (2) over the system's update-block partials.
__device__ inline void reduce_direction(const double* __restrict__ cur_p, const double* __restrict__ prev_p,
                                        const double* __restrict__ bb_p, int stride, int b0, int cnt,
                                        double (&cur)[4], double (&prev)[4], double (&bb)[2]) {
  double res[10];
  block_sums<10>(cnt, [&](int t, double (&v)[10]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[q] += cur_p[q * stride + b0 + t];
      v[4 + q] += prev_p[q * stride + b0 + t];
    }
    v[8] += bb_p[b0 + t];
    v[9] += bb_p[stride + b0 + t];
  }, res);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    cur[q] = res[q];
    prev[q] = res[4 + q];
  }
  bb[0] = res[8];
  bb[1] = res[9];
}

// The update kernel's prologue: residual dots (4) and |b|^2 (2) over the
// system's update blocks, p.Ap (2) over its direction blocks.
__device__ inline void reduce_update(const double* __restrict__ cur_p, const double* __restrict__ bb_p, int su,
                                     int u0, int nu_cnt, const double* __restrict__ pap_p, int sd, int d0, int nd_cnt,
                                     double (&cur)[4], double (&bb)[2], double (&pap)[2]) {
  double res[8];
  block_sums<8>(nu_cnt > nd_cnt ? nu_cnt : nd_cnt, [&](int t, double (&v)[8]) {
    if (t < nu_cnt) {
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] += cur_p[q * su + u0 + t];
      v[4] += bb_p[u0 + t];
      v[5] += bb_p[su + u0 + t];
    }
    if (t < nd_cnt) {
      v[6] += pap_p[d0 + t];
      v[7] += pap_p[sd + d0 + t];
    }
  }, res);
#pragma unroll
  for (int q = 0; q < 4; ++q) cur[q] = res[q];
  bb[0] = res[4];
  bb[1] = res[5];
  pap[0] = res[6];
  pap[1] = res[7];
}

// Block partial of NV values into part[k * stride + blockIdx.x].
template <int NV>
__device__ inline void write_partials(double (&vals)[NV], double* __restrict__ part, int stride) {
  xr_block_partials<NV>(vals, part, stride);
}

// Per right-hand side: has this system's residual reached its current target?
// At a (re)start the residual is the true one and the target is the final
// tolerance; in between, mixed precision aims for the inner drop.
__device__ inline void rhs_converged(const double (&cur)[4], const double (&bb)[2], const CgSys& c, int j,
                                     bool mixed, double tol2, bool (&conv)[2]) {
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    double target = tol2 * bb[r];
    if (mixed && j > 0) target = fmax(target, INNER_DROP2 * c.rr0[r]);
    conv[r] = cur[2 + r] <= target;
  }
}

__device__ inline void settle(CgSys* c, CgCounters* cnt, int state, int k) {
  if (state == CG_REFRESH) c->kr = k + 1;  // the next direction launch refreshes
  c->done = state;
  if (state != CG_REFRESH) atomicAdd(&cnt->settled, 1);
}

// Block-Jacobi preconditioner.  Consecutive atoms joined by a strong
// same-image bond (raw order >= 0.3 to one of the previous three atoms of the
// block) form blocks of at most four atoms, e.g. one water molecule; blocks
// never cross an update block, so they depend only on the system's own atoms.
// M^-1 is the inverse of the blocks of H on its diagonal (image terms
// included).  Every update block holds whole preconditioner blocks, so
// applying M^-1 needs only the block's residuals, exchanged through LDS.
struct BlockRows {
  int first[UPD_BLOCK];  // local row of each row's block start
  int size[UPD_BLOCK];
};

// Preconditioner blocks of an update block's rows [0, rows).  First the
// bonded runs, as one pass in row order would find them: row t joins the
// current run when it is fewer than RUN_MAX rows past the run's start and
// bonded to an earlier row of it.  A row with no strong bond to the three
// rows before it always starts a run, so each such row's thread walks on to
// the next one and all walks run in parallel.  Then whole consecutive runs
// are packed greedily into blocks of at most BLOCK_MAX rows (thread 0, over
// the run starts held as a bit mask); a block row finds its block start and
// size from the mask of block starts.
__device__ inline void assign_blocks(const unsigned char* __restrict__ join_sh, int rows, BlockRows& br) {
  __shared__ unsigned long long run_mask[UPD_BLOCK / 64], blk_mask[UPD_BLOCK / 64];
  const int t = threadIdx.x;
  auto starts = [&](int r) { return r == 0 || (join_sh[r] & 7u) == 0; };
  if (t < rows && starts(t)) {
    int start = t;
    br.first[t] = t;
    for (int r = t + 1; r < rows && !starts(r); ++r) {
      bool joins = false;
      if (r - start < RUN_MAX) {
        for (int d = 1; d <= 3; ++d)
          if (((join_sh[r] >> (d - 1)) & 1u) && r - d >= start) joins = true;
      }
      if (!joins) start = r;
      br.first[r] = start;
    }
  }
  __syncthreads();
  const unsigned long long m = __ballot(t < rows && br.first[t] == t);
  if ((t & 63) == 0) run_mask[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    // Runs start at s_0 = 0 < s_1 < ...; run k ends where run k+1 starts (or
    // at rows).  A run that would take the current block past BLOCK_MAX rows
    // opens a new block (runs are at most RUN_MAX <= BLOCK_MAX rows).
    blk_mask[0] = 1ull;
    for (int w = 1; w < UPD_BLOCK / 64; ++w) blk_mask[w] = 0ull;
    int bstart = 0, run = 0;
    for (int w = 0; w < UPD_BLOCK / 64; ++w) {
      unsigned long long bits = run_mask[w];
      while (bits) {
        const int next = w * 64 + __ffsll((long long)bits) - 1;
        bits &= bits - 1;
        if (next == 0) continue;
        if (next - bstart > BLOCK_MAX) {
          bstart = run;
          blk_mask[run >> 6] |= 1ull << (run & 63);
        }
        run = next;
      }
    }
    if (rows - bstart > BLOCK_MAX) blk_mask[run >> 6] |= 1ull << (run & 63);
  }
  __syncthreads();
  if (t < rows) {
    // block start: the highest block-start bit at or below t; end: the next one above t (or rows)
    int f = 0;
    for (int w = t >> 6; w >= 0; --w) {
      unsigned long long bits = blk_mask[w];
      if (w == (t >> 6)) bits &= (t & 63) == 63 ? ~0ull : ((1ull << ((t & 63) + 1)) - 1);
      if (bits) {
        f = w * 64 + 63 - __clzll((long long)bits);
        break;
      }
    }
    int e = rows;
    for (int w = t >> 6; w < UPD_BLOCK / 64; ++w) {
      unsigned long long bits = blk_mask[w];
      if (w == (t >> 6)) bits &= (t & 63) == 63 ? 0ull : ~((1ull << ((t & 63) + 1)) - 1);
      if (bits) {
        e = w * 64 + __ffsll((long long)bits) - 1;
        break;
      }
    }
    br.first[t] = f;
    br.size[t] = (e < rows ? e : rows) - f;
  }
  __syncthreads();
}


__global__ void __launch_bounds__(UPD_BLOCK) k_cg_init(Systems S, const int* __restrict__ type, FFDev ff,
                                                       const double2* __restrict__ rhs,
                                                       const double2* __restrict__ guess, double2* __restrict__ X,
                                                       double2* __restrict__ R, double4* __restrict__ ZP,
                                                       float4* __restrict__ ZQ, const int* __restrict__ row_ptr,
                                                       const int* __restrict__ nbr_j, const double* __restrict__ hval,
                                                       const unsigned char* __restrict__ join,
                                                       double* __restrict__ minv, int2* __restrict__ binfo,
                                                       double* __restrict__ eta_out, double* __restrict__ part_b,
                                                       double* __restrict__ part_bb, int nu, CgSys* __restrict__ cg,
                                                       CgCounters* __restrict__ cnt,
                                                       const unsigned* __restrict__ abort_flag, int n_all) {
  const int sys = block_system(S, blockIdx.x, false);
  const SysDev& sd = S.sys[sys];
  const int local = blockIdx.x - sd.upd0;
  const int row0 = sd.atom0 + local * UPD_BLOCK;
  const int rows = min(UPD_BLOCK, sd.atom0 + sd.n - row0);
  const int t = threadIdx.x;
  const int i = row0 + t;
  const bool mine = t < rows;
  __shared__ unsigned char join_sh[UPD_BLOCK];
  __shared__ BlockRows br;
  __shared__ double hb[UPD_BLOCK][BLOCK_MAX];  // each row's row of its H block, inverted in place
  __shared__ double2 bsh[UPD_BLOCK];
  join_sh[t] = mine ? join[i] : 0;
  __syncthreads();
  assign_blocks(join_sh, rows, br);
  double eta = 0.0;
  double2 b = make_double2(0.0, 0.0);
  const int f = mine ? br.first[t] : t, m = mine ? br.size[t] : 0;
  if (mine) {
    const double* a = ff.atom + type[i] * AP_COUNT;
    eta = a[AP_ETA];
    b = rhs ? rhs[i] : make_double2(-a[AP_CHI], -1.0);
    bsh[t] = b;
    // row t of its block of H: eta on the diagonal plus h summed over all
    // images to each block mate, from the row's neighbors in [g0, g0 + m)
#pragma unroll
    for (int c = 0; c < BLOCK_MAX; ++c) hb[t][c] = c == t - f ? eta : 0.0;
    const int g0 = row0 + f;
    int lo = row_ptr[i], hi = row_ptr[i + 1];
    const int e_end = hi;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (nbr_j[mid] < g0) lo = mid + 1;
      else hi = mid;
    }
    for (int e = lo; e < e_end; ++e) {
      const int jj = nbr_j[e];
      if (jj >= g0 + m) break;
      hb[t][jj - g0] += hval[e];
    }
  }
  __syncthreads();
  // In-place Gauss-Jordan inversion of every block, one thread per block row
  // (no pivoting: the blocks of H are symmetric positive definite).  Each
  // step reads the old pivot row before any row of the block is rewritten.
  for (int p = 0; p < BLOCK_MAX; ++p) {
    const bool act = mine && p < m;
    double prow[BLOCK_MAX];
    double pp = 1.0, fr = 0.0;
    if (act) {
#pragma unroll
      for (int c = 0; c < BLOCK_MAX; ++c) prow[c] = hb[f + p][c];
      pp = hb[f + p][p];
      fr = hb[t][p];
    }
    __syncthreads();
    if (act) {
      const double piv = 1.0 / pp;
      if (t == f + p) {
#pragma unroll
        for (int c = 0; c < BLOCK_MAX; ++c) hb[t][c] = (c == p ? 1.0 : prow[c]) * piv;
      } else {
#pragma unroll
        for (int c = 0; c < BLOCK_MAX; ++c) hb[t][c] = (c == p ? 0.0 : hb[t][c]) - fr * ((c == p ? 1.0 : prow[c]) * piv);
      }
    }
    __syncthreads();
  }
  double vals[4] = {0, 0, 0, 0};
  double bb[2] = {0, 0};
  if (mine) {
    double2 z = make_double2(0.0, 0.0);
    for (int c = 0; c < m; ++c) {
      z.x += hb[t][c] * bsh[f + c].x;
      z.y += hb[t][c] * bsh[f + c].y;
      minv[(size_t)c * n_all + i] = hb[t][c];
    }
    binfo[i] = make_int2(f, m);
    X[i] = guess ? guess[i] : make_double2(0.0, 0.0);
    R[i] = b;
    ZP[i] = make_double4(z.x, z.y, 0.0, 0.0);  // (z_s, z_t, p_s, p_t), p = 0 before iteration 0
    if (ZQ) ZQ[i] = make_float4((float)z.x, (float)z.y, 0.0f, 0.0f);
    eta_out[i] = eta;
    vals[0] = b.x * z.x;
    vals[1] = b.y * z.y;
    vals[2] = b.x * b.x;
    vals[3] = b.y * b.y;
    bb[0] = vals[2];
    bb[1] = vals[3];
  }
  // slot 1 holds the residual state before iteration 0
  write_partials<4>(vals, part_b + 4 * nu, nu);
  write_partials<2>(bb, part_bb, nu);
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt->abort = abort_flag ? (int)(*abort_flag != 0) : 0;
  if (local == 0 && threadIdx.x == 0) {
    CgSys c{};
    c.done = CG_RUNNING;
    c.iters = 0;
    c.k0 = 0;
    c.restarts = 0;
    cg[sys] = c;
    if (S.err[sys]) {
      // an input error was flagged upstream: nothing to solve
      cg[sys].done = CG_FAILED;
      atomicAdd(&cnt->settled, 1);
    } else if (guess) {
      // warm start: launch 0 computes the true residual of the guess
      cg[sys].done = CG_REFRESH;
      cg[sys].kr = 0;
      cg[sys].k0 = -1;
    }
  }
}

// z = M^-1 r for one row of an update block, from the block's residuals in LDS
// (bi = local row of the block start, block size).
// A row's block-Jacobi inverse row (column-major storage: entry c of row i at
// minv[c * stride + i]), loaded into registers in one unrolled batch so the
// loads are in flight together.
struct MinvRow {
  double w[BLOCK_MAX];
  int2 bi;
};

__device__ inline void load_minv(const double* __restrict__ minv, int stride, int i, bool mine,
                                 const int2* __restrict__ binfo, MinvRow& m) {
  m.bi = mine ? binfo[i] : make_int2(0, 0);
#pragma unroll
  for (int c = 0; c < BLOCK_MAX; ++c) m.w[c] = c < m.bi.y ? minv[(size_t)c * stride + i] : 0.0;
}

// z = M^-1 r for one row from the block's residuals in LDS
__device__ inline double2 apply_block(const MinvRow& m, const double2* __restrict__ rsh) {
  double2 z = make_double2(0.0, 0.0);
#pragma unroll
  for (int c = 0; c < BLOCK_MAX; ++c) {
    if (c < m.bi.y) {
      const double2 r = rsh[m.bi.x + c];
      z.x += m.w[c] * r.x;
      z.y += m.w[c] * r.y;
    }
  }
  return z;
}

// Eight H entries per lane of one row (coalesced over the wave): columns and
// values, padded with the row atom and zero.
template <class HV>
struct RowBatch {
  int j[8];
  HV h[8];
};

template <class HV>
__device__ inline void load_batch(const int* __restrict__ nbr_j, const HV* __restrict__ hv, int base, int e1, int i,
                                  int lane, RowBatch<HV>& b) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int e = base + u * 64 + lane;
    const bool ok = e < e1;
    b.j[u] = ok ? nbr_j[e] : i;
    b.h[u] = ok ? hv[e] : HV(0);
  }
}

// sum_e h_e p_j for one batch, p_j = z_j + beta p_old_j from the gathered records
__device__ inline void gather_batch(const RowBatch<double>& b, const double4* __restrict__ zp, double bs, double bt,
                                    double& as, double& at) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const double4 v = zp[b.j[u]];
    as += b.h[u] * (v.x + bs * v.z);
    at += b.h[u] * (v.y + bt * v.w);
  }
}

__device__ inline void gather_batch(const RowBatch<float>& b, const float4* __restrict__ zq, float bs, float bt,
                                    float& as, float& at) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const float4 v = zq[b.j[u]];
    as += b.h[u] * (v.x + bs * v.z);
    at += b.h[u] * (v.y + bt * v.w);
  }
}

// The same sums with z and p_old kept apart (a = sum h z, sum h p_old per
// right-hand side), for the row gathered before beta is known.
__device__ inline void gather_split(const RowBatch<double>& b, const double4* __restrict__ zp, double (&a)[4]) {
  // two groups of four 32-B gathers: all eight in flight at once would not
  // fit the 128 registers of a 1024-thread block next to the prefetched row
#pragma unroll
  for (int g = 0; g < 8; g += 4) {
#pragma unroll
    for (int u = g; u < g + 4; ++u) {
      const double4 v = zp[b.j[u]];
      a[0] += b.h[u] * v.x;
      a[1] += b.h[u] * v.y;
      a[2] += b.h[u] * v.z;
      a[3] += b.h[u] * v.w;
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

__device__ inline void gather_split(const RowBatch<float>& b, const float4* __restrict__ zq, float (&a)[4]) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const float4 v = zq[b.j[u]];
    a[0] += b.h[u] * v.x;
    a[1] += b.h[u] * v.y;
    a[2] += b.h[u] * v.z;
    a[3] += b.h[u] * v.w;
  }
}

// Iteration k, first half: beta, p = z + beta p_old and A p (rows of the
// system strided over its blocks, one wave per row), partial p.Ap per block.
// zp_in holds (z, p_old) of every atom; the new p goes to the p half of
// zp_out.  Rows are software-pipelined: each wave loads the first 512 H
// entries of its next row before gathering for the current one, and those of
// its first row before the prologue reduction, so the H stream overlaps the
// gathers.
template <bool MIXED>
__global__ void __launch_bounds__(DIR_BLOCK) k_cg_direction(Systems S, int k, const int* __restrict__ row_ptr,
                                                            const int* __restrict__ nbr_j,
                                                            const double* __restrict__ hval,
                                                            const float* __restrict__ hval32,
                                                            const double* __restrict__ eta,
                                                            const double4* __restrict__ zp_in,
                                                            double4* __restrict__ zp_out,
                                                            const float4* __restrict__ zq_in,
                                                            float4* __restrict__ zq_out, double2* __restrict__ AP,
                                                            const double* __restrict__ part_b,
                                                            const double* __restrict__ part_bb, int nu,
                                                            double* __restrict__ part_a, int nd, double tol2,
                                                            int max_iter, CgSys* __restrict__ cg,
                                                            CgCounters* __restrict__ cnt, const int* __restrict__ type,
                                                            FFDev ff, const double2* __restrict__ rhs,
                                                            const double2* __restrict__ X, double2* __restrict__ R) {
  const int sys = block_system(S, blockIdx.x, true);
  const CgSys c = cg[sys];
  const SysDev& sd = S.sys[sys];
  const int local = blockIdx.x - sd.dir0;
  const int lane = threadIdx.x & 63;
  const int waves = sd.ndir * (DIR_BLOCK >> 6);
  const int end = sd.atom0 + sd.n;
  if (c.done == CG_REFRESH && c.kr == k) {
    // Refresh: FP64 true residual r = b - A x.  The update launch applies the
    // preconditioner, writes (z, p = 0) and reduces r.z and r.r.
    for (int i = sd.atom0 + local * (DIR_BLOCK >> 6) + (threadIdx.x >> 6); i < end; i += waves) {
      double as = 0.0, at = 0.0;
      for (int e = row_ptr[i] + lane; e < row_ptr[i + 1]; e += 64) {
        const double2 v = X[nbr_j[e]];
        as += hval[e] * v.x;
        at += hval[e] * v.y;
      }
      as = xr_wave_sum(as);
      at = xr_wave_sum(at);
      if (lane == 0) {
        const double* a = ff.atom + type[i] * AP_COUNT;
        const double2 b = rhs ? rhs[i] : make_double2(-a[AP_CHI], -1.0);
        const double2 xv = X[i];
        const double d = eta[i];
        R[i] = make_double2(b.x - (d * xv.x + as), b.y - (d * xv.y + at));
      }
    }
    return;
  }
  // A system refreshed by the previous launch (k0 == k, set by its update
  // kernel) runs again from j = 0; one in any other state sits this launch out.
  const bool resumed = c.done == CG_REFRESH && c.k0 == k;
  if (c.done != CG_RUNNING && !resumed) return;
  const int j = k - c.k0;
  using HV = typename std::conditional<MIXED, float, double>::type;
  const HV* hv;
  if constexpr (MIXED) hv = hval32;
  else hv = hval;
  int i = sd.atom0 + local * (DIR_BLOCK >> 6) + (threadIdx.x >> 6);
  int e0 = 0, e1 = 0;
  RowBatch<HV> cur_b;
  if (i < end) {
    e0 = row_ptr[i];
    e1 = row_ptr[i + 1];
    load_batch(nbr_j, hv, e0, e1, i, lane, cur_b);
  }
  // The wave's first row is gathered before beta is known (sum h z and
  // sum h p_old apart), so the prologue reduction overlaps it; the next row's
  // first 512 entries are already on their way.
  int inext = i + waves, n0 = 0, n1 = 0;
  RowBatch<HV> next_b;
  if (inext < end) {
    n0 = row_ptr[inext];
    n1 = row_ptr[inext + 1];
    load_batch(nbr_j, hv, n0, n1, inext, lane, next_b);
  }
  double first[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < end) {
    if constexpr (MIXED) {
      float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      gather_split(cur_b, zq_in, f);
      for (int base = e0 + 512; base < e1; base += 512) {
        RowBatch<HV> more;
        load_batch(nbr_j, hv, base, e1, i, lane, more);
        gather_split(more, zq_in, f);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) first[q] = (double)f[q];
    } else {
      gather_split(cur_b, zp_in, first);
      for (int base = e0 + 512; base < e1; base += 512) {
        RowBatch<HV> more;
        load_batch(nbr_j, hv, base, e1, i, lane, more);
        gather_split(more, zp_in, first);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) first[q] = xr_wave_sum(first[q]);
  double cur[4], prev[4], bb[2];
  reduce_direction(part_b + ((k - 1) & 1) * 4 * nu, part_b + (k & 1) * 4 * nu, part_bb, nu, sd.upd0, sd.nupd, cur,
                   prev, bb);
  bool conv[2];
  rhs_converged(cur, bb, c, j, MIXED, tol2, conv);
  if ((conv[0] && conv[1]) || c.iters >= max_iter || (MIXED && j >= INNER_MAX)) {
    if (local == 0 && threadIdx.x == 0) {
      int state;
      if (conv[0] && conv[1] && (!MIXED || j == 0)) state = CG_DONE;
      else if (c.iters >= max_iter || c.restarts > MAX_RESTARTS) state = CG_STALLED;
      else state = CG_REFRESH;
      settle(&cg[sys], cnt, state, k);
    }
    return;
  }
  if (j == 0 && local == 0 && threadIdx.x == 0) {
    if (MIXED) {
      cg[sys].rr0[0] = cur[2];
      cg[sys].rr0[1] = cur[3];
    }
    if (resumed) cg[sys].done = CG_RUNNING;
  }
  double beta_s = 0.0, beta_t = 0.0;
  if (j > 0) {
    if (!conv[0] && prev[0] != 0.0) beta_s = cur[0] / prev[0];
    if (!conv[1] && prev[1] != 0.0) beta_t = cur[1] / prev[1];
  }
  double pap_s = 0.0, pap_t = 0.0;
  // p = z + beta p_old and A p = eta p + sum h p_j for row i from its sums
  auto finish_row = [&](int row, double acc_s, double acc_t) {
    const double4 v = zp_in[row];
    const double2 pr = make_double2(v.x + beta_s * v.z, v.y + beta_t * v.w);
    const double d = eta[row];
    const double2 ap = make_double2(d * pr.x + acc_s, d * pr.y + acc_t);
    reinterpret_cast<double2*>(zp_out + row)[1] = pr;
    if (MIXED) reinterpret_cast<float2*>(zq_out + row)[1] = make_float2((float)pr.x, (float)pr.y);
    AP[row] = ap;
    pap_s += pr.x * ap.x;
    pap_t += pr.y * ap.y;
  };
  if (i < end && lane == 0) finish_row(i, first[0] + beta_s * first[2], first[1] + beta_t * first[3]);
  i = inext;
  e0 = n0;
  e1 = n1;
  cur_b = next_b;
  while (i < end) {
    inext = i + waves;
    n0 = 0;
    n1 = 0;
    if (inext < end) {
      n0 = row_ptr[inext];
      n1 = row_ptr[inext + 1];
      load_batch(nbr_j, hv, n0, n1, inext, lane, next_b);
    }
    double acc_s, acc_t;
    if constexpr (MIXED) {
      const float bs = (float)beta_s, bt = (float)beta_t;
      float fs = 0.0f, ft = 0.0f;
      gather_batch(cur_b, zq_in, bs, bt, fs, ft);
      for (int base = e0 + 512; base < e1; base += 512) {
        RowBatch<HV> more;
        load_batch(nbr_j, hv, base, e1, i, lane, more);
        gather_batch(more, zq_in, bs, bt, fs, ft);
      }
      acc_s = (double)fs;
      acc_t = (double)ft;
    } else {
      acc_s = 0.0;
      acc_t = 0.0;
      gather_batch(cur_b, zp_in, beta_s, beta_t, acc_s, acc_t);
      for (int base = e0 + 512; base < e1; base += 512) {
        RowBatch<HV> more;
        load_batch(nbr_j, hv, base, e1, i, lane, more);
        gather_batch(more, zp_in, beta_s, beta_t, acc_s, acc_t);
      }
    }
    acc_s = xr_wave_sum(acc_s);
    acc_t = xr_wave_sum(acc_t);
    if (lane == 0) finish_row(i, acc_s, acc_t);
    i = inext;
    e0 = n0;
    e1 = n1;
    cur_b = next_b;
  }
  double vals[2] = {pap_s, pap_t};
  write_partials<2>(vals, part_a, nd);
}

// Iteration k, second half: alpha, x += alpha p, r -= alpha A p, z = M^-1 r
// (into the z half of zp), partial r.z and r.r per block.
template <bool MIXED>
__global__ void __launch_bounds__(UPD_BLOCK) k_cg_update(Systems S, int k, int n_all, const double* __restrict__ minv,
                                                         const int2* __restrict__ binfo,
                                                         double2* __restrict__ X, double2* __restrict__ R,
                                                         double4* __restrict__ zp, float4* __restrict__ zq,
                                                         const double2* __restrict__ AP,
                                                         const double* __restrict__ part_a, int nd,
                                                         double* __restrict__ part_b,
                                                         const double* __restrict__ part_bb, int nu, double tol2,
                                                         CgSys* __restrict__ cg) {
  const int sys = block_system(S, blockIdx.x, false);
  const CgSys c = cg[sys];
  const SysDev& sd = S.sys[sys];
  const int local = blockIdx.x - sd.upd0;
  __shared__ double2 rsh[UPD_BLOCK];
  const int i = sd.atom0 + local * UPD_BLOCK + threadIdx.x;
  const bool mine = i < sd.atom0 + sd.n;
  if (c.done != CG_RUNNING && !(c.done == CG_REFRESH && c.kr == k)) return;
  MinvRow mrow;
  load_minv(minv, n_all, i, mine, binfo, mrow);
  if (c.done == CG_REFRESH) {
    // A refresh requested by this iteration's direction launch waits for the
    // next one; a true residual computed by this iteration's direction launch
    // is preconditioned and reduced here.
    if (c.kr != k) return;
    double vals[4] = {0, 0, 0, 0};
    const double2 r = mine ? R[i] : make_double2(0.0, 0.0);
    rsh[threadIdx.x] = r;
    __syncthreads();
    if (mine) {
      const double2 z = apply_block(mrow, rsh);
      zp[i] = make_double4(z.x, z.y, 0.0, 0.0);
      if (MIXED) zq[i] = make_float4((float)z.x, (float)z.y, 0.0f, 0.0f);
      vals[0] = r.x * z.x;
      vals[1] = r.y * z.y;
      vals[2] = r.x * r.x;
      vals[3] = r.y * r.y;
    }
    write_partials<4>(vals, part_b + (k & 1) * 4 * nu, nu);
    if (local == 0 && threadIdx.x == 0) {
      cg[sys].k0 = k + 1;
      cg[sys].restarts = c.restarts + 1;
    }
    return;
  }
  if (c.done != CG_RUNNING) return;
  double cur[4], pap[2], bb[2];
  reduce_update(part_b + ((k - 1) & 1) * 4 * nu, part_bb, nu, sd.upd0, sd.nupd, part_a, nd, sd.dir0, sd.ndir, cur,
                bb, pap);
  bool conv[2];
  rhs_converged(cur, bb, c, k - c.k0, MIXED, tol2, conv);
  const double alpha_s = (!conv[0] && pap[0] != 0.0) ? cur[0] / pap[0] : 0.0;
  const double alpha_t = (!conv[1] && pap[1] != 0.0) ? cur[1] / pap[1] : 0.0;
  double vals[4] = {0, 0, 0, 0};
  double2 r = make_double2(0.0, 0.0);
  if (mine) {
    const double2 p = reinterpret_cast<const double2*>(zp + i)[1];
    const double2 ap = AP[i];
    double2 xv = X[i];
    r = R[i];
    xv.x += alpha_s * p.x;
    xv.y += alpha_t * p.y;
    r.x -= alpha_s * ap.x;
    r.y -= alpha_t * ap.y;
    X[i] = xv;
    R[i] = r;
  }
  rsh[threadIdx.x] = r;
  __syncthreads();
  if (mine) {
    const double2 z = apply_block(mrow, rsh);
    reinterpret_cast<double2*>(zp + i)[0] = z;
    if (MIXED) reinterpret_cast<float2*>(zq + i)[0] = make_float2((float)z.x, (float)z.y);
    vals[0] = r.x * z.x;
    vals[1] = r.y * z.y;
    vals[2] = r.x * r.x;
    vals[3] = r.y * r.y;
  }
  write_partials<4>(vals, part_b + (k & 1) * 4 * nu, nu);
  if (local == 0 && threadIdx.x == 0) cg[sys].iters = c.iters + 1;
}

// Per system: q = s + mu t with mu = (Q - sum s) / sum t; the (s, t) solution
// is kept in `keep` (when given) as the next call's starting guess.
__global__ void __launch_bounds__(1024) k_qeq_charges(Systems S, const double2* __restrict__ X,
                                                      double* __restrict__ q, double* __restrict__ mu_out,
                                                      double2* __restrict__ keep, int neutral) {
  const SysDev& sd = S.sys[blockIdx.x];
  __shared__ double red[2][16];
  double ss = 0.0, st = 0.0;
  for (int i = sd.atom0 + threadIdx.x; i < sd.atom0 + sd.n; i += blockDim.x) {
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
    mu_s = ((neutral ? 0.0 : sd.total_charge) - a) / b;
    if (mu_out) mu_out[blockIdx.x] = mu_s;
  }
  __syncthreads();
  const double mu = mu_s;
  for (int i = sd.atom0 + threadIdx.x; i < sd.atom0 + sd.n; i += blockDim.x) {
    const double2 v = X[i];
    q[i] = v.x + mu * v.y;
    if (keep) keep[i] = v;
  }
}

}  // namespace

hipError_t xr_edge_prep(int n, const double* x, const int* type, const Systems& S, FFDev ff, const int* row_ptr,
                        const int* nbr_j, const int* nbr_s, double bond_cut, double* hval, float* hval32,
                        unsigned char* bflag, int* bond_count, unsigned char* join, hipStream_t s) {
  const int blocks = (n + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;
  XR_LAUNCH(XT_EDGE, k_edge_prep, dim3(blocks), dim3(256), 0, s, n, x, type, S, ff, row_ptr, nbr_j, nbr_s,
            bond_cut, hval, hval32, bflag, bond_count, join);
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

// Direction blocks per system: about DIR_ROWS_PER_WAVE rows per wave up to a cap (XREAC_CG_MAX_BLOCKS
// overrides the cap for tuning; it must be the same for every call that is
// compared bitwise, since a system's blocks depend only on its size and this cap).
void xr_cg_blocks(int n, int* ndir, int* nupd) {
  static const int cap = [] {
    const char* v = std::getenv("XREAC_CG_MAX_BLOCKS");
    const int c = v ? std::atoi(v) : DIR_MAX_BLOCKS;
    return c > 0 ? c : DIR_MAX_BLOCKS;
  }();
  const int rows_per_block = DIR_ROWS_PER_WAVE * (DIR_BLOCK / 64);
  int d = (n + rows_per_block - 1) / rows_per_block;
  *ndir = d < 1 ? 1 : (d > cap ? cap : d);
  const int u = (n + UPD_BLOCK - 1) / UPD_BLOCK;
  *nupd = u < 1 ? 1 : u;
}

size_t xr_cg_partials(int nd, int nu) { return (size_t)(2 * 4 * nu + 2 * nu + 2 * nd) + 64; }

size_t xr_cg_block_doubles(int n) { return (size_t)n * BLOCK_MAX; }

size_t xr_cg_state_bytes(int ns) { return sizeof(CgSys) * (size_t)ns + sizeof(CgCounters) + 64; }

hipError_t xr_cg_solve(const CgArgs& a, hipStream_t s, CgOutcome* out) {
  const int nu = a.nu, nd = a.nd, ns = a.S.ns;
  double* part_b = a.partials;
  double* part_bb = part_b + 2 * 4 * nu;
  double* part_a = part_bb + 2 * nu;
  CgCounters* cnt = static_cast<CgCounters*>(a.state);
  CgSys* cg = xr_cg_systems(a.state);
  const double tol2 = a.tol * a.tol;
  const bool mixed = a.hval32 != nullptr;
  float4* zq0 = mixed ? a.ZQ0 : nullptr;
  float4* zq1 = mixed ? a.ZQ1 : nullptr;
  // counters start at zero (cleared with the control block)
  hipLaunchKernelGGL(k_cg_init, dim3(nu), dim3(UPD_BLOCK), 0, s, a.S, a.type, a.ff, a.rhs, a.guess, a.X, a.R, a.ZP0,
                     zq0, a.row_ptr, a.nbr_j, a.hval, a.join, a.minv, a.binfo, a.eta, part_b, part_bb, nu, cg, cnt,
                     a.abort_flag, a.n);
  XR_LAUNCH_CHECK();
  int k = 0;
  // Every system settles within max_iter executed iterations plus two
  // launches per refresh (the request and the FP64 sweep; one more for a warm start).
  const long long k_limit = (long long)a.max_iter + 3LL * (MAX_RESTARTS + 3) + a.first_check + a.check_every;
  out->aborted = false;
  out->exhausted = false;
  auto enqueue = [&](int iters) {
    for (int c = 0; c < iters && k < k_limit; ++c) {
      double4* zp_in = (k & 1) ? a.ZP1 : a.ZP0;
      double4* zp_out = (k & 1) ? a.ZP0 : a.ZP1;
      float4* zq_in = (k & 1) ? zq1 : zq0;
      float4* zq_out = (k & 1) ? zq0 : zq1;
      if (mixed) {
        XR_LAUNCH(XT_CG_DIRECTION, k_cg_direction<true>, dim3(nd), dim3(DIR_BLOCK), 0, s, a.S, k, a.row_ptr,
                  a.nbr_j, a.hval, a.hval32, a.eta, zp_in, zp_out, zq_in, zq_out, a.AP, part_b, part_bb, nu, part_a,
                  nd, tol2, a.max_iter, cg, cnt, a.type, a.ff, a.rhs, a.X, a.R);
        XR_LAUNCH(XT_CG_UPDATE, k_cg_update<true>, dim3(nu), dim3(UPD_BLOCK), 0, s, a.S, k, a.n, a.minv, a.binfo, a.X, a.R,
                  zp_out, zq_out, a.AP, part_a, nd, part_b, part_bb, nu, tol2, cg);
      } else {
        XR_LAUNCH(XT_CG_DIRECTION, k_cg_direction<false>, dim3(nd), dim3(DIR_BLOCK), 0, s, a.S, k, a.row_ptr,
                  a.nbr_j, a.hval, a.hval32, a.eta, zp_in, zp_out, zq_in, zq_out, a.AP, part_b, part_bb, nu, part_a,
                  nd, tol2, a.max_iter, cg, cnt, a.type, a.ff, a.rhs, a.X, a.R);
        XR_LAUNCH(XT_CG_UPDATE, k_cg_update<false>, dim3(nu), dim3(UPD_BLOCK), 0, s, a.S, k, a.n, a.minv, a.binfo, a.X, a.R,
                  zp_out, zq_out, a.AP, part_a, nd, part_b, part_bb, nu, tol2, cg);
      }
      ++k;
    }
    return hipGetLastError();
  };
  // Read-back of (settled, abort) into a pinned slot behind each chunk.  The
  // first chunk is sized from the previous solve and waited for at once; if
  // systems are still running, later chunks are issued one ahead of the slot
  // being waited for, so the device keeps working while the host reads.
  // Launches that run after every system has settled find nothing to do.
  auto post = [&](int slot) {
    hipError_t e = hipMemcpyAsync(a.poll_host + 2 * slot, cnt, 2 * sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(a.poll_ev[slot], s);
    return e;
  };
  bool pending[2] = {false, false};
  auto issue = [&](int slot, int iters) {
    if (k >= k_limit) return hipSuccess;
    hipError_t e = enqueue(iters);
    if (e == hipSuccess) e = post(slot);
    if (e == hipSuccess) pending[slot] = true;
    return e;
  };
  hipError_t e = issue(0, a.first_check);
  if (e != hipSuccess) return e;
  int cur = 0;
  bool pipelined = false;
  for (;;) {
    e = hipEventSynchronize(a.poll_ev[cur]);
    if (e != hipSuccess) return e;
    pending[cur] = false;
    const int* st = a.poll_host + 2 * cur;
    if (st[1]) {  // buffers overflowed upstream; the caller repeats the evaluation
      out->aborted = true;
      return hipSuccess;
    }
    if (st[0] >= ns) break;
    if (!pipelined) {  // from here on one chunk runs while the host reads the other's slot
      pipelined = true;
      e = issue(1 - cur, a.check_every);
      if (e != hipSuccess) return e;
    }
    e = issue(cur, a.check_every);
    if (e != hipSuccess) return e;
    if (pending[1 - cur]) {
      cur = 1 - cur;
    } else if (!pending[cur]) {
      out->exhausted = true;
      break;
    }
  }
  out->sweeps = k;
  hipLaunchKernelGGL(k_qeq_charges, dim3(ns), dim3(1024), 0, s, a.S, a.X, a.q, a.mu, a.keep, a.neutral ? 1 : 0);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

