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
// edges (one 16-byte gather of (m_s, m_t) per edge).
//
// The solver is the pipelined preconditioned CG of Ghysels and Vanroose
// (2014): its only matrix product per iteration, n = A m, reads a vector the
// previous iteration finished, and its two dot products come from the same
// pass, so one CG iteration is one kernel.  Each block re-reduces the previous
// launch's per-block partials in a fixed order, so no separate reduction
// launches are needed and the result is deterministic.
//
// Batches: every system owns a fixed set of blocks and chunks (sized from its
// atom count only), reduces only its own partials and stops on its own
// convergence test, so a system's charges are bitwise the same alone and in
// any batch.
//
// Mixed precision: the A m sweep reads FP32 copies of H and of the gathered m
// while the recurrence vectors and every dot product stay FP64.  When the
// recurrence residual has dropped by the inner factor, the system's next
// launch replaces it by the FP64 true residual r = b - A x (one FP64 sweep),
// the launch after it computes w = A u in FP64, and the recurrence restarts
// from there (iterative refinement); a system is converged only when a true
// residual meets the tolerance.  A warm start begins with the same refresh.
// All of it runs on the device: the host only polls how many systems have
// settled.
#include <algorithm>
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
constexpr int BOND_QUEUE = 128;     // edge pass: per-wave queue of edges within the bond cutoff

// ---------------------------------------------------------------- edge pass
// With a slab (the fill's sorted rows), the pass first moves each row's keys
// to their CSR places, nbr_j[e] and nbr_s[e]; without one (a supplied list)
// it reads them there.  The row's H-bond candidates (a hydrogen of H-bond
// type 1 and an acceptor of type 2 within min(7.5, rc), either way round)
// are compacted, in row order, as (j, shift code) pairs into the first
// hb_count[i] slots of the row's own range of hb_cand, so the H-bond terms
// read only them instead of rescanning the row.
__global__ void __launch_bounds__(256) k_edge_prep(int n, const double* __restrict__ x, const int* __restrict__ type,
                                                   Systems S, FFDev ff, const int* __restrict__ row_ptr,
                                                   const unsigned long long* __restrict__ slab, int slab_stride,
                                                   int* __restrict__ nbr_j, int* __restrict__ nbr_s,
                                                   unsigned short* __restrict__ col16, double bond_cut, double* __restrict__ hval,
                                                   float* __restrict__ hval32, unsigned char* __restrict__ bflag,
                                                   int* __restrict__ bond_count, unsigned char* __restrict__ join,
                                                   int2* __restrict__ hb_cand, int* __restrict__ hb_count) {
  const int i = blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const SysDev& sd = S.sys[__builtin_amdgcn_readfirstlane(S.asys[i])];
  const CellDev& cell = sd.cell;
  const int ti = type[i];
  const double rc = ff.gen[12];
  const double cut = 0.01 * ff.gen[29];
  const double hcut = fmin(7.5, rc);
  const double hb_i = ff.atom[ti * AP_COUNT + AP_HBOND];
  int cnt = 0;
  unsigned jm = 0;  // bit d-1: strong same-image bond to atom i - d (d = 1..3), for the CG blocks
  const int zero = xr_pack_shift(0, 0, 0);
  const int e_begin = row_ptr[i], e_end = row_ptr[i + 1];
  const unsigned long long* srow = slab ? slab + (size_t)i * slab_stride - e_begin : nullptr;
  // Edges within the bond cutoff (about one in eight) wait in a wave-private
  // LDS queue and have their raw bond order evaluated 64 at a time, so the
  // exp-heavy code runs on full waves instead of masked ones.
  const int w = threadIdx.x >> 6;
  __shared__ int q_e[ROWS_PER_BLOCK][BOND_QUEUE], q_j[ROWS_PER_BLOCK][BOND_QUEUE], q_t[ROWS_PER_BLOCK][BOND_QUEUE];
  __shared__ double q_r[ROWS_PER_BLOCK][BOND_QUEUE];
  __shared__ unsigned char q_f[ROWS_PER_BLOCK][BOND_QUEUE];  // 0x80: zero image shift
  int qn = 0;  // identical in every lane
  int hn = 0;  // H-bond candidates so far, identical in every lane
  auto flush = [&](int count) {  // the first count (<= 64) queued edges, one per lane
    if (lane < count) {
      const int e = q_e[w][lane], j = q_j[w][lane];
      const double r = q_r[w][lane];
      const unsigned char f = q_f[w][lane];
      BondRaw b;
      xr_bond_raw(ff, ti, q_t[w][lane], r, b);
      const double raw = (b.sig + b.pi) + b.pp;
      const bool active = raw >= cut;
      if (raw >= BLOCK_BOND && j < i && j >= i - 3 && (f & 0x80)) jm |= 1u << (i - j - 1);
      bflag[e] = active ? EF_BOND : 0;
      cnt += active;
    }
  };
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int base = e_begin; base < e_end; base += 64) {
    const int e = base + lane;
    const bool valid = e < e_end;
    bool cand = false, hb = false;
    int j = 0, tj = 0, code = 0;
    double r = 0.0;
    if (valid) {
      if (srow) {
        const unsigned long long key = srow[e];
        j = (int)(key >> 32);
        code = (int)(key & 0xffffffffu);
        nbr_j[e] = j;
        nbr_s[e] = code;
      } else {
        j = nbr_j[e];
        code = nbr_s[e];
      }
      if (col16) col16[e] = (unsigned short)(j - sd.atom0);  // system-relative
      tj = type[j];
      double v[3];
      r = xr_edge(x, cell, i, j, code, v);
      double s, ds;
      xr_shield(r, rc, ff.pair[(ti * ff.nt + tj) * PP_COUNT + PP_GAMMA], s, ds);
      hval[e] = 14.4 * s;
      if (hval32) hval32[e] = (float)(14.4 * s);
      if (r <= hcut) {
        const double hb_j = ff.atom[tj * AP_COUNT + AP_HBOND];
        hb = (hb_i == 1.0 && hb_j == 2.0) || (hb_i == 2.0 && hb_j == 1.0);
      }
      cand = r <= bond_cut;
      if (!cand) bflag[e] = 0;
    }
    const unsigned long long hm = __ballot(hb);
    if (hb) hb_cand[e_begin + hn + __popcll(hm & below)] = make_int2(j, code);
    hn += __popcll(hm);
    const unsigned long long cm = __ballot(cand);
    if (cand) {
      const int slot = qn + __popcll(cm & below);
      q_e[w][slot] = e;
      q_j[w][slot] = j;
      q_t[w][slot] = tj;
      q_r[w][slot] = r;
      q_f[w][slot] = (unsigned char)(code == zero ? 0x80 : 0);
    }
    qn += __popcll(cm);
    if (qn >= 64) {
      xr_wave_sync();
      flush(64);
      xr_wave_sync();
      // queue entries past the batch move to the front (reads from 64.., writes below 64)
      if (lane < qn - 64) {
        q_e[w][lane] = q_e[w][64 + lane];
        q_j[w][lane] = q_j[w][64 + lane];
        q_t[w][lane] = q_t[w][64 + lane];
        q_r[w][lane] = q_r[w][64 + lane];
        q_f[w][lane] = q_f[w][64 + lane];
      }
      xr_wave_sync();
      qn -= 64;
    }
  }
  xr_wave_sync();
  flush(qn);
  cnt = xr_wave_sum_int(cnt);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) jm |= __shfl_down(jm, off, 64);
  if (lane == 0) {
    bond_count[i] = cnt;
    join[i] = (unsigned char)jm;
    hb_count[i] = hn;
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
// Kernel boundaries order all cross-block communication: each launch writes
// per-block partial dot products, and every block of the next launch reduces
// its system's partials again in the same fixed order (redundant but cheap,
// and bitwise reproducible).  Device-scope fences inside a kernel are avoided
// on purpose: on a multi-XCD part they write back the whole L2.
// Two block shapes, chosen by a system's own size (so a system is solved the
// same way alone and in any batch): systems of CG_SMALL_ATOMS atoms and more
// take 1024-thread blocks over 48-row chunks, which fill the chip from
// 12,288 atoms on; smaller systems take 512-thread (or, in a batch of many,
// 256-thread) blocks over 24-row chunks, which spread a system over twice as
// many CUs and pack a batch of small systems more tightly.  Each iteration
// launches one kernel per shape present in the batch; small systems' blocks
// come first.
constexpr int CG_BLOCK = 1024;            // 16 waves; the rows of a chunk are spread over the waves
constexpr int CG_MAX_BLOCKS = 256;        // per system: few partial sums to reduce
constexpr int CHUNK = 48;                 // nominal rows per chunk; chunks start at bonded-run starts
constexpr int CG_BLOCK_SMALL = 512;       // 8 waves
// A batch with many small-system blocks launches them as 256-thread blocks
// instead (same chunks and arithmetic, so bitwise the same results): under a
// concurrent flood of one-wave workgroups from the side streams, a 4-wave
// block finds room on a CU far sooner than an 8-wave one, while a lone small
// system keeps the 8-wave shape (fewer rows per wave, shorter latency chain).
constexpr int CG_BLOCK_NARROW = 256;
constexpr int CG_NARROW_MIN_BLOCKS = 1024;
constexpr int CG_MAX_BLOCKS_SMALL = 512;
constexpr int CHUNK_SMALL = 24;
constexpr int CG_SMALL_ATOMS = 8192;
constexpr int CHUNK_MAX = CG_CHUNK_ROWS_MAX;  // CHUNK + RUN_MAX - 1 rows at most: one wave holds a chunk
constexpr int INNER_MAX = 64;          // mixed precision: iterations between FP64 restarts
constexpr double INNER_DROP2 = 1e-8;   // mixed precision: squared residual drop before a restart
constexpr int MAX_RESTARTS = 12;
constexpr int NDOT = 6;                // per block: r.u, w.u, r.r for both right-hand sides
constexpr int CG_LEAD_LAUNCHES = 4;    // launches queued before the caller's side-stream work
static_assert(CHUNK + RUN_MAX - 1 <= CHUNK_MAX && CHUNK_SMALL + RUN_MAX - 1 <= CHUNK_MAX,
              "a chunk must fit one wave");

// System owning global chunk b: chunk offsets (upd0) ascend with the system
// index.  (Iteration-block offsets do not: small systems' blocks come first.)
__device__ inline int chunk_system(const Systems& S, int b) {
  int lo = 0, hi = S.ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int start = S.sys[mid].upd0;
    if (start <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Local row where chunk c of a system starts: the nominal c * sd.chunk, moved
// forward to the first row within RUN_MAX rows that starts a bonded run (no
// strong bond to any of the three rows before it), so that molecules are not
// split between chunks.  Called by a whole wave.
__device__ inline int chunk_start(const SysDev& sd, const unsigned char* __restrict__ join, int c) {
  if (c <= 0) return 0;
  const int t = c * sd.chunk;
  if (t >= sd.n) return sd.n;
  const int lane = threadIdx.x & 63;
  const int r = t + lane;
  const bool ok = lane < RUN_MAX && r < sd.n && (join[sd.atom0 + r] & 7u) == 0;
  const unsigned long long m = __ballot(ok);
  return m ? t + __ffsll((long long)m) - 1 : t;
}

// Block-Jacobi preconditioner.  Consecutive atoms joined by a strong
// same-image bond (raw order >= 0.3 to one of the previous three atoms of the
// run) form bonded runs of at most RUN_MAX atoms (molecules); whole runs are
// packed into blocks of at most BLOCK_MAX atoms that never cross a chunk.
// M^-1 is the inverse of the blocks of H on its diagonal (image terms
// included).  A chunk is processed by one block of the iteration kernel, so
// applying M^-1 needs only the chunk's own rows, exchanged through LDS.
struct ChunkBlocks {
  int first[CHUNK_MAX];  // local row of each row's block start
  int size[CHUNK_MAX];
};

// Preconditioner blocks of a chunk's rows [0, rows), computed by one wave
// (lane t owns row t).  First the bonded runs, as one pass in row order would
// find them: row t joins the current run when it is fewer than RUN_MAX rows
// past the run's start and bonded to an earlier row of it.  A row with no
// strong bond to the three rows before it always starts a run, so each such
// row's lane walks on to the next one and all walks run in parallel.  Then
// whole consecutive runs are packed greedily into blocks of at most BLOCK_MAX
// rows (the same scalar loop in every lane, over the run starts as a bit
// mask); a row finds its block start and size from the mask of block starts.
__device__ inline void assign_blocks(const unsigned char* __restrict__ join_sh, int rows, ChunkBlocks& br) {
  const int t = threadIdx.x & 63;
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
  xr_wave_sync();
  const unsigned long long runs = __ballot(t < rows && br.first[t] == t);
  // Runs start at s_0 = 0 < s_1 < ...; run k ends where run k+1 starts (or
  // at rows).  A run that would take the current block past BLOCK_MAX rows
  // opens a new block (runs are at most RUN_MAX <= BLOCK_MAX rows).
  unsigned long long blk = 1ull, bits = runs & ~1ull;
  int bstart = 0, run = 0;
  while (bits) {
    const int next = __ffsll((long long)bits) - 1;
    bits &= bits - 1;
    if (next - bstart > BLOCK_MAX) {
      bstart = run;
      blk |= 1ull << run;
    }
    run = next;
  }
  if (rows - bstart > BLOCK_MAX) blk |= 1ull << run;
  if (t < rows) {
    const unsigned long long upto = t == 63 ? ~0ull : ((2ull << t) - 1);
    const int f = 63 - __clzll((long long)(blk & upto));
    const unsigned long long above = blk & ~upto;
    const int e = above ? __ffsll((long long)above) - 1 : rows;
    br.first[t] = f;
    br.size[t] = (e < rows ? e : rows) - f;
  }
  xr_wave_sync();
}

// Right-hand sides of atom i: the supplied pair, else (-chi_i, -1).
__device__ inline double2 cg_rhs(const double2* __restrict__ rhs, const int* __restrict__ type, const FFDev& ff,
                                 int i) {
  if (rhs) return rhs[i];
  return make_double2(-ff.atom[type[i] * AP_COUNT + AP_CHI], -1.0);
}

// One block per chunk, one lane per row: chunk bounds, the preconditioner
// blocks and their inverses, x0 and, for a cold start, r0 = b and u0 = M^-1 b.
// The first iteration launch then computes w0 = A u0 (a warm start first
// computes r0 = b - A x0 and u0 there) and the solve state starts as a
// refresh that is already under way.
__global__ void __launch_bounds__(64) k_cg_init(Systems S, const int* __restrict__ type, FFDev ff,
                                                const double2* __restrict__ rhs, const double2* __restrict__ guess,
                                                CgVec v, const int* __restrict__ row_ptr,
                                                const int* __restrict__ nbr_j, const double* __restrict__ hval,
                                                const unsigned char* __restrict__ join, double* __restrict__ minv,
                                                int2* __restrict__ binfo, int2* __restrict__ chunks,
                                                CgBlock* __restrict__ blocks, double* __restrict__ eta_out,
                                                CgSys* __restrict__ cg,
                                                CgCounters* __restrict__ cnt, const unsigned* __restrict__ abort_flag,
                                                int n_all) {
  const int sys = chunk_system(S, blockIdx.x);
  const SysDev& sd = S.sys[sys];
  const int c = blockIdx.x - sd.upd0;
  const int r0 = chunk_start(sd, join, c);
  const int r1 = chunk_start(sd, join, c + 1);
  const int rows = r1 - r0;
  const int t = threadIdx.x;
  const int row0 = sd.atom0 + r0;
  const int i = row0 + t;
  const bool mine = t < rows;
  __shared__ unsigned char join_sh[CHUNK_MAX];
  __shared__ ChunkBlocks br;
  __shared__ double hb[CHUNK_MAX][BLOCK_MAX];  // each row's row of its H block, inverted in place
  __shared__ double2 bsh[CHUNK_MAX];
  if (t == 0) chunks[blockIdx.x] = make_int2(row0, sd.atom0 + r1);
  {
    // the iteration block owning this chunk first: its system, this chunk's
    // rows, the system-local index of this chunk and the rows' pointers
    const int per = (sd.nupd + sd.ndir - 1) / sd.ndir;
    if (c % per == 0) {
      CgBlock& blk = blocks[sd.dir0 + c / per];
      if (t == 0) blk.desc = make_int4(sys, row0, sd.atom0 + r1, c);
      blk.rp[t] = row_ptr[row0 + min(t, rows)];
      if (t == 0) blk.rp[CHUNK_MAX] = row_ptr[row0 + rows];
    }
  }
  join_sh[t] = mine ? join[i] : 0;
  xr_wave_sync();
  assign_blocks(join_sh, rows, br);
  double eta = 0.0;
  double2 b = make_double2(0.0, 0.0);
  const int f = mine ? br.first[t] : t, m = mine ? br.size[t] : 0;
  if (mine) {
    eta = ff.atom[type[i] * AP_COUNT + AP_ETA];
    b = cg_rhs(rhs, type, ff, i);
    bsh[t] = b;
    // row t of its block of H: eta on the diagonal plus h summed over all
    // images to each block mate, from the row's neighbors in [g0, g0 + m)
#pragma unroll
    for (int q = 0; q < BLOCK_MAX; ++q) hb[t][q] = q == t - f ? eta : 0.0;
    const int g0 = row0 + f;
    int lo = row_ptr[i], hi = row_ptr[i + 1];
    const int e_end = hi;
    // First entry with j >= g0 (rows are sorted by j): an 8-way search whose
    // seven probes per round are loaded together, then the last few entries.
    while (hi - lo > 8) {
      const int step = (hi - lo) >> 3;
      int below = 0;
#pragma unroll
      for (int k = 0; k < 7; ++k) below += nbr_j[lo + (k + 1) * step] < g0;
      const int nhi = below < 7 ? lo + (below + 1) * step : hi;
      lo = below > 0 ? lo + below * step + 1 : lo;
      hi = nhi;
    }
    while (lo < hi && nbr_j[lo] < g0) ++lo;
    // The block's entries follow: a window of 16 loaded together, then one by one.
    constexpr int WIN = 16;
    int wj[WIN];
    double wh[WIN];
#pragma unroll
    for (int k = 0; k < WIN; ++k) {
      const bool ok = lo + k < e_end;
      wj[k] = ok ? nbr_j[lo + k] : 0x7fffffff;
      wh[k] = ok ? hval[lo + k] : 0.0;
    }
    bool more = true;
#pragma unroll
    for (int k = 0; k < WIN; ++k) {
      if (more && wj[k] < g0 + m) hb[t][wj[k] - g0] += wh[k];
      else more = false;
    }
    if (more) {
      for (int e = lo + WIN; e < e_end; ++e) {
        const int jj = nbr_j[e];
        if (jj >= g0 + m) break;
        hb[t][jj - g0] += hval[e];
      }
    }
  }
  xr_wave_sync();
  // In-place Gauss-Jordan inversion of every block, one lane per block row
  // (no pivoting: the blocks of H are symmetric positive definite).  Each
  // step reads the old pivot row before any row of the block is rewritten.
  for (int p = 0; p < BLOCK_MAX; ++p) {
    const bool act = mine && p < m;
    double prow[BLOCK_MAX];
    double pp = 1.0, fr = 0.0;
    if (act) {
#pragma unroll
      for (int q = 0; q < BLOCK_MAX; ++q) prow[q] = hb[f + p][q];
      pp = hb[f + p][p];
      fr = hb[t][p];
    }
    xr_wave_sync();
    if (act) {
      const double piv = 1.0 / pp;
      if (t == f + p) {
#pragma unroll
        for (int q = 0; q < BLOCK_MAX; ++q) hb[t][q] = (q == p ? 1.0 : prow[q]) * piv;
      } else {
#pragma unroll
        for (int q = 0; q < BLOCK_MAX; ++q) hb[t][q] = (q == p ? 0.0 : hb[t][q]) - fr * ((q == p ? 1.0 : prow[q]) * piv);
      }
    }
    xr_wave_sync();
  }
  if (mine) {
    double2 u = make_double2(0.0, 0.0);
    for (int q = 0; q < m; ++q) {
      u.x += hb[t][q] * bsh[f + q].x;
      u.y += hb[t][q] * bsh[f + q].y;
      minv[(size_t)q * n_all + i] = hb[t][q];
    }
    binfo[i] = make_int2(f, m);
    eta_out[i] = eta;
    v.X[i] = guess ? guess[i] : make_double2(0.0, 0.0);
    if (!guess) {
      v.R[i] = b;
      v.U[i] = u;
    }
  }
  if (blockIdx.x == 0 && t == 0) cnt->abort = abort_flag ? (int)(*abort_flag != 0) : 0;
  if (c == 0 && t == 0) {
    CgCtl s{};
    s.done = CG_REFRESH;
    s.kr = guess ? 0 : -1;  // launch kr computes r and u, launch kr + 1 computes w
    CgSys init{};
    if (S.err[sys]) {
      // an input error was flagged upstream: nothing to solve
      s.done = CG_FAILED;
      atomicAdd(&cnt->settled, 1);
    }
    init.ctl[0] = s;
    init.ctl[1] = s;
    cg[sys] = init;
  }
}

// The sweep n = A m over a chunk's rows [r0, r1): wave w owns rows r0 + w,
// r0 + w + WAVES, ...; lanes 2k and 2k + 1 of the wave hold the entry range
// [row_ptr[i], row_ptr[i + 1]) of its k-th row, loaded for all of them at
// once.  Every load below is unconditional, with the index clamped to a valid
// entry, and the lanes past a row's end are masked where the values are used,
// so the compiler keeps a batch's loads in flight together instead of waiting
// behind per-lane branches.
__device__ inline int wave_index() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

template <int WAVES>
__device__ inline int wave_rows(int r0, int r1) {
  const int first = r0 + wave_index();
  return first < r1 ? (r1 - first + WAVES - 1) / WAVES : 0;
}

template <int WAVES>
__device__ inline int wave_row_ptrs(int r0, int r1, const int* __restrict__ row_ptr) {
  const int lane = threadIdx.x & 63;
  const int i = r0 + wave_index() + (lane >> 1) * WAVES + (lane & 1);
  return row_ptr[min(i, r1)];  // lanes past the wave's rows read a valid pointer, never used
}

// U entries per lane of one row, coalesced over the wave: entry
// base + 64 u + lane in slot u, as loaded (slots at or past `end` hold the
// row's last entry, or entry 0 for an empty row, and are masked by the user).
// A lane sums its entries in entry order whatever U is, so U changes no result.
template <class HV, int U>
struct Batch {
  unsigned j[U];
  HV h[U];
};

// Entries per lane and batch: 4 in every shape.  The smaller shapes then
// hold 5-6 waves per SIMD instead of 4; the 1024-thread block (85 VGPRs in
// FP64 against 125 with 8 entries) leaves room on its CU for the side
// stream's waves, so it no longer waits for a whole CU to drain.
template <int BLOCK>
constexpr int batch_entries() {
  return 4;
}

// Columns are the 4-byte neighbor indices or, in batches of at most 65,536
// atoms, their 2-byte copies relative to the system's first atom
// (IT = unsigned short).  Slot u reads entry min(base + 64 u + lane, last)
// with last = max(end - 1, base): for an empty row that is entry `base`,
// which lies inside the edge arrays (they are carved with at least one spare
// entry past the list).  The addresses are the row's first entry (uniform)
// plus a 32-bit lane offset.
template <class HV, class IT, int U>
__device__ inline void load_batch(const IT* __restrict__ cols, const HV* __restrict__ hv, int base, int end,
                                  Batch<HV, U>& b) {
  const int lane = threadIdx.x & 63;
  const unsigned last = (unsigned)max(end - 1 - base, 0);
  const char* cb = reinterpret_cast<const char*>(cols + base);
  const char* hb = reinterpret_cast<const char*>(hv + base);
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const unsigned r = min((unsigned)(u * 64 + lane), last);
    b.j[u] = (unsigned)*reinterpret_cast<const IT*>(cb + r * (unsigned)sizeof(IT));
    b.h[u] = *reinterpret_cast<const HV*>(hb + r * (unsigned)sizeof(HV));
  }
}

// m_j for column j: a 32-bit byte offset from the system's base, so the
// gather addresses stay one register each
template <class VT>
__device__ inline VT gather(const char* __restrict__ vbase, unsigned j) {
  return *reinterpret_cast<const VT*>(vbase + j * (unsigned)sizeof(VT));
}

// Off-diagonal row sums sum_e h_e m_j(e) of the wave's rows.  `cur` holds the
// first batch of the wave's first row (loaded by the caller).  The wave works
// through its rows' batches in order; each step issues the gathers of the
// current batch (masked slots gather the system's first atom with value 0),
// then the loads of the next batch, then accumulates, so the next batch is in
// flight while the gathers return.  Per lane the products are summed in entry
// order, then over the wave; lane 0 leaves each row's pair of sums in
// out[row - r0].
template <int WAVES, class HV, class VT, class AT, class IT, int U>
__device__ inline void chunk_spmv(int r0, int r1, int rp, const IT* __restrict__ cols, const HV* __restrict__ hv,
                                  const VT* __restrict__ vec, int off, Batch<HV, U>& cur, double2* out) {
  const int lane = threadIdx.x & 63;
  const int nk = wave_rows<WAVES>(r0, r1);
  if (nk == 0) return;
  const int first = wave_index();  // the wave's first row, relative to r0
  // columns are relative to off (the system's first atom for 2-byte columns, else 0)
  const char* vbase = reinterpret_cast<const char*>(vec + off);
  int k = 0;
  int base = __builtin_amdgcn_readlane(rp, 0), end = __builtin_amdgcn_readlane(rp, 1);
  AT as = AT(0), at = AT(0);
  for (;;) {
    // the next batch: the rest of this row, else the next row's first (an
    // empty range once the wave's rows are done)
    int kn = k, nbase = base + 64 * U, nend = end;
    if (nbase >= end) {
      kn = k + 1;
      const int l = kn < nk ? 2 * kn : 0;
      nbase = __builtin_amdgcn_readlane(rp, l);
      nend = kn < nk ? __builtin_amdgcn_readlane(rp, l + 1) : nbase;
    }
    VT g[U];
#pragma unroll
    for (int u = 0; u < U; ++u) g[u] = gather<VT>(vbase, base + u * 64 + lane < end ? cur.j[u] : 0u);
    Batch<HV, U> nxt;
    load_batch(cols, hv, nbase, nend, nxt);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const HV h = base + u * 64 + lane < end ? cur.h[u] : HV(0);
      as += h * g[u].x;
      at += h * g[u].y;
    }
    if (kn != k) {
      const double s0 = xr_wave_sum((double)as), s1 = xr_wave_sum((double)at);
      if (lane == 0) out[first + k * WAVES] = make_double2(s0, s1);
      if (kn >= nk) return;
      as = AT(0);
      at = AT(0);
    }
    cur = nxt;
    k = kn;
    base = nbase;
    end = nend;
  }
}

// The first batch of the wave's first row of a chunk (an empty range for a
// wave without rows)
template <int WAVES, class HV, class IT, int U>
__device__ inline void first_batch(int r0, int r1, int rp, const IT* __restrict__ cols, const HV* __restrict__ hv,
                                   Batch<HV, U>& b) {
  const int base = __builtin_amdgcn_readlane(rp, 0);
  const int end = wave_rows<WAVES>(r0, r1) > 0 ? __builtin_amdgcn_readlane(rp, 1) : base;
  load_batch(cols, hv, base, end, b);
}

// Sum over the series of a system's ndir block partials in p[0..ndir): lane l
// adds partials l, l + 64, ... in that order (the loads issued together).
template <int NP>
__device__ inline double partial_sum(const double* __restrict__ p, int ndir) {
  const int l = threadIdx.x & 63;
  double v[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) v[q] = p[min(l + 64 * q, ndir - 1)];
  double a = l < ndir ? v[0] : 0.0;
#pragma unroll
  for (int q = 1; q < NP; ++q) a = l + 64 * q < ndir ? a + v[q] : a;
  for (int t = l + 64 * NP; t < ndir; t += 64) a += p[t];  // block caps above 64 NP
  return a;
}

// What a system's block does in launch k, from the state the previous launch
// left (j: iterations since the last (re)start).
enum CgMode { MODE_IDLE = 0, MODE_RESIDUAL = 1, MODE_START = 2, MODE_ITER = 3 };

__device__ inline int cg_mode(const CgCtl& c, int k, int& j) {
  j = 0;
  if (c.done == CG_RUNNING) {
    j = k - c.k0;
    return MODE_ITER;
  }
  if (c.done == CG_REFRESH) {
    if (k == c.kr) return MODE_RESIDUAL;
    if (k == c.kr + 1) return MODE_START;
    if (k == c.kr + 2) return MODE_ITER;
  }
  return MODE_IDLE;
}

// Row operands staged through LDS: each thread loads one vector entry and one
// inverse-block column before the chunk's sweep and stores them after it.
enum CgItem { IT_Z = 0, IT_Q, IT_S, IT_P, IT_X, IT_R, IT_U, IT_W, IT_M, IT_B, NITEM };

struct ChunkLds {
  double2 st[NITEM][CHUNK_MAX];
  double mv[BLOCK_MAX][CHUNK_MAX];
  double eta[CHUNK_MAX];
  int2 bi[CHUNK_MAX];
  double2 sums[CHUNK_MAX];   // off-diagonal A v row sums
  double ex[2][CHUNK_MAX];   // vector exchanged for the M^-1 apply, per right-hand side
};

// Items staged per row: NITEM vectors, eta, the block placement and
// BLOCK_MAX inverse-block columns, item w + u WAVES by wave w in pass u.
constexpr int STAGE_ITEMS = 16;
static_assert(NITEM + 2 <= STAGE_ITEMS && BLOCK_MAX <= STAGE_ITEMS, "staged items");

// Per pass: one vector entry (items below NITEM), eta (item NITEM, in val.x)
// or the block placement (item NITEM + 1, its bits in val.x), and one
// inverse-block column entry.
template <int WAVES>
struct Staged {
  static_assert(STAGE_ITEMS % WAVES == 0, "whole passes");
  double2 val[STAGE_ITEMS / WAVES];
  double mv[STAGE_ITEMS / WAVES];
};

__device__ inline bool item_used(int mode, int item) {
  if (mode == MODE_ITER) return item <= IT_M;
  if (mode == MODE_RESIDUAL) return item == IT_X || item == IT_B;
  return item == IT_R || item == IT_U || item == IT_B;  // MODE_START
}

template <int WAVES>
__device__ inline void stage_write(int r0, int r1, const Staged<WAVES>& s, ChunkLds& L) {
  const int row = threadIdx.x & 63;
  if (r0 + row >= r1) return;
#pragma unroll
  for (int u = 0; u < STAGE_ITEMS / WAVES; ++u) {
    const int item = (threadIdx.x >> 6) + u * WAVES;
    L.mv[item][row] = s.mv[u];
    if (item < NITEM) L.st[item][row] = s.val[u];
    if (item == NITEM) L.eta[row] = s.val[u].x;
    if (item == NITEM + 1) *reinterpret_cast<long long*>(&L.bi[row]) = __double_as_longlong(s.val[u].x);
  }
}

// The 1024-thread shape (one item per thread) keeps the staged values in
// registers over the sweep and writes them to LDS after it, so their loads
// finish under the gathers; the smaller shapes (two or four items per
// thread) write them at once, which costs a wait before the sweep but keeps
// their registers for the sweep.
template <int WAVES>
constexpr bool stage_held() { return STAGE_ITEMS / WAVES == 1; }

template <int WAVES>
__device__ inline void stage_load(int mode, int r0, int r1, const CgVec& v, const double2* __restrict__ m_in,
                                  const double* __restrict__ minv, int n_all, const int2* __restrict__ binfo,
                                  const double* __restrict__ eta, const double2* __restrict__ rhs,
                                  const int* __restrict__ type, const FFDev& ff, Staged<WAVES>& s, ChunkLds& L) {
  // lane t stages row r0 + t; lanes past the chunk load row r1 - 1 (never stored)
  const int i = min(r0 + (int)(threadIdx.x & 63), r1 - 1);
#pragma unroll
  for (int u = 0; u < STAGE_ITEMS / WAVES; ++u) {
    const int item = (threadIdx.x >> 6) + u * WAVES;  // uniform over the wave
    double2 val = make_double2(0.0, 0.0);
    const double mv = minv[(size_t)item * n_all + i];  // entries past the block size are never used
    if (item < NITEM && item_used(mode, item)) {
      if (item == IT_B) {
        val = cg_rhs(rhs, type, ff, i);
      } else {
        const double2* src = item == IT_Z ? v.Z : item == IT_Q ? v.Q : item == IT_S ? v.S : item == IT_P ? v.P
                           : item == IT_X ? v.X : item == IT_R ? v.R : item == IT_U ? v.U : item == IT_W ? v.W : m_in;
        val = src[i];
      }
    }
    if (item == NITEM) val.x = eta[i];
    if (item == NITEM + 1) val.x = __longlong_as_double(*reinterpret_cast<const long long*>(binfo + i));
    s.val[u] = val;
    s.mv[u] = mv;
  }
  if constexpr (!stage_held<WAVES>()) stage_write<WAVES>(r0, r1, s, L);
}

// After the sweep: the held values into LDS (the smaller shapes wrote theirs before it)
template <int WAVES>
__device__ inline void stage_store(int r0, int r1, const Staged<WAVES>& s, ChunkLds& L) {
  if constexpr (stage_held<WAVES>()) stage_write<WAVES>(r0, r1, s, L);
}


// M^-1 applied to the exchanged vector of right-hand side h for row t
__device__ inline double apply_block(const ChunkLds& L, int h, int t) {
  const int2 bi = L.bi[t];
  double z = 0.0;
#pragma unroll
  for (int c = 0; c < BLOCK_MAX; ++c)
    if (c < bi.y) z += L.mv[c][t] * L.ex[h][bi.x + c];
  return z;
}

struct CgStep {
  double alpha[2], beta[2];
  bool fresh;  // first iteration since a (re)start: the old z, q, s, p are not read
};

// The chunk's rows after its sweep, waves 0 and 1 (lane t owns row t, wave h
// right-hand side h):
//   iteration: n = A m;  z = n + b z, q = m + b q, s = w + b s, p = u + b p;
//              x += a p, r -= a s, u -= a q, w -= a z;  m = M^-1 w
//   residual:  r = b - A x;  u = M^-1 r
//   start:     w = A u;  m = M^-1 w
// and the wave's dots (r.u, w.u, r.r and, at a start, b.b; after a residual
// only r.r and b.b) accumulate in dots.
template <bool MIXED>
__device__ inline void chunk_tail(int mode, int r0, int r1, const CgStep& st, const CgVec& v, double2* m_out,
                                  float2* m32_out, ChunkLds& L, double (&dots)[4]) {
  const int t = threadIdx.x & 63;
  const int h = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar store addresses
  const int i = r0 + t;
  const bool mine = i < r1;
  auto get = [&](int item) { return h ? L.st[item][t].y : L.st[item][t].x; };
  auto put = [&](double2* vec, double val) { reinterpret_cast<double*>(vec + i)[h] = val; };
  double r = 0.0, u = 0.0, w = 0.0;
  if (mine) {
    const double eta = L.eta[t];
    const double ax = h ? L.sums[t].y : L.sums[t].x;
    if (mode == MODE_ITER) {
      const double a = h ? st.alpha[1] : st.alpha[0];
      const double b = h ? st.beta[1] : st.beta[0];
      const double m = get(IT_M);
      const double nn = eta * m + ax;
      u = get(IT_U);
      w = get(IT_W);
      r = get(IT_R);
      double x = get(IT_X);
      double z = nn, q = m, s = w, p = u;
      if (!st.fresh) {
        z = nn + b * get(IT_Z);
        q = m + b * get(IT_Q);
        s = w + b * get(IT_S);
        p = u + b * get(IT_P);
      }
      x += a * p;
      r -= a * s;
      u -= a * q;
      w -= a * z;
      put(v.Z, z);
      put(v.Q, q);
      put(v.S, s);
      put(v.P, p);
      put(v.X, x);
      put(v.R, r);
      put(v.U, u);
      put(v.W, w);
      L.ex[h][t] = w;
    } else if (mode == MODE_RESIDUAL) {
      r = get(IT_B) - (eta * get(IT_X) + ax);
      put(v.R, r);
      L.ex[h][t] = r;
    } else {
      r = get(IT_R);
      u = get(IT_U);
      w = eta * u + ax;
      put(v.W, w);
      L.ex[h][t] = w;
    }
  }
  xr_wave_sync();
  if (!mine) return;
  const double mz = apply_block(L, h, t);
  if (mode == MODE_RESIDUAL) {
    put(v.U, mz);
    const double b = get(IT_B);
    dots[2] += r * r;  // the true residual's r.r and b.b, for the start launch's convergence test
    dots[3] += b * b;
    return;
  }
  put(m_out, mz);
  if (MIXED) reinterpret_cast<float*>(m32_out + i)[h] = (float)mz;
  dots[0] += r * u;
  dots[1] += w * u;
  dots[2] += r * r;
  if (mode == MODE_START) {
    const double b = get(IT_B);
    dots[3] += b * b;
  }
}

__device__ inline void settle(CgSys* c, CgCounters* cnt, const CgCtl& cur, int state, int k) {
  CgCtl s = cur;
  s.done = state;
  if (state == CG_REFRESH) {
    s.kr = k + 1;  // the next launch computes the true residual
    c->ctl[k & 1] = s;
    return;
  }
  // A final state is written to both slots: every block of this launch stops
  // whichever slot it reads.
  c->ctl[0] = s;
  c->ctl[1] = s;
  c->iters = cur.it;
  c->ks = k;
  atomicAdd(&cnt->settled, 1);
}

// Launch k of the pipelined preconditioned CG (Ghysels and Vanroose, 2014),
// two right-hand sides.  With gamma = r.u, delta = w.u from the previous
// launch, beta = gamma / gamma_old and alpha = gamma / (delta - beta gamma /
// alpha_old) (alpha = gamma / delta at a start), every launch does one sweep
// n = A m over its rows and the vector recurrences listed at chunk_tail.  The
// sweep reads only m, computed for every row by the previous launch, so one
// launch per iteration suffices; the dots for the next launch come out of the
// same pass.  Each block of a system owns whole chunks.  The launch covers
// iteration blocks [base, base + gridDim.x), all of one shape (BLOCK threads,
// at most 128 VGPRs).
template <bool MIXED, int BLOCK, class IT>
__global__ void __launch_bounds__(BLOCK, 4) k_cg_iteration(Systems S, int k, CgVec v, const int* __restrict__ row_ptr,
                                                           const IT* __restrict__ nbr_j,
                                                           const double* __restrict__ hval,
                                                           const float* __restrict__ hval32,
                                                           const double* __restrict__ eta,
                                                           const double* __restrict__ minv,
                                                           const int2* __restrict__ binfo,
                                                           const int2* __restrict__ chunks,
                                                           const CgBlock* __restrict__ blocks,
                                                           const int* __restrict__ type, FFDev ff,
                                                           const double2* __restrict__ rhs, double* __restrict__ part,
                                                           double* __restrict__ part_bb, int nd, double tol2,
                                                           int max_iter, CgSys* __restrict__ cg,
                                                           CgCounters* __restrict__ cnt, int n_all, int base) {
  constexpr int WAVES = BLOCK / 64;
  constexpr int QUARTER = BLOCK / 4;  // threads per series pair of partials
  static_assert(WAVES % 4 == 0, "a quarter of the block per series pair");
  // partial loads per lane in flight together: the block cap of the systems this shape runs
  constexpr int NP = (BLOCK == CG_BLOCK ? CG_MAX_BLOCKS : CG_MAX_BLOCKS_SMALL) / 64;
  const int b = base + blockIdx.x;
  // The block's descriptor gives its system and first chunk, and its record
  // holds that chunk's row pointers, so the waves' first loads do not wait on
  // each other.
  const int4 desc = blocks[b].desc;
  const int sys = desc.x;
  // the wave's first-chunk row pointers: loaded by block index alongside the descriptor
  int rp = blocks[b].rp[min(wave_index() + (int)((threadIdx.x & 63) >> 1) * WAVES + (int)(threadIdx.x & 1), CHUNK_MAX)];
  const CgCtl c = cg[sys].ctl[(k - 1) & 1];
  int j;
  const int mode = cg_mode(c, k, j);
  if (mode == MODE_IDLE) return;
  const SysDev& sd = S.sys[sys];
  const int coff = std::is_same<IT, unsigned short>::value ? sd.atom0 : 0;  // 2-byte columns are system-relative
  const int local = b - sd.dir0;
  const int per = (sd.nupd + sd.ndir - 1) / sd.ndir;
  const int c0 = desc.w, c1 = min(c0 + per, sd.nupd);
  __shared__ ChunkLds L;
  const double2* m_in = v.M[k & 1];
  double2* m_out = v.M[(k + 1) & 1];
  float2* m32_out = v.M32[(k + 1) & 1];
  const bool lead = local == 0 && threadIdx.x == 0;
  double dots[4] = {0, 0, 0, 0};

  CgStep st{};
  if (mode == MODE_ITER) {
    using HV = typename std::conditional<MIXED, float, double>::type;
    using VT = typename std::conditional<MIXED, float2, double2>::type;
    using AT = HV;
    const HV* hv;
    const VT* vin;
    if constexpr (MIXED) {
      hv = hval32;
      vin = v.M32[k & 1];
    } else {
      hv = hval;
      vin = m_in;
    }
    // The first chunk's first rows, its row operands and the previous
    // launch's partials are all loaded before the sweep; the partials are
    // reduced after it (the sweep n = A m needs neither alpha nor beta), so
    // their latency hides under the gathers.  Each quarter of the block holds
    // one series pair (in order r.u, w.u, r.r, b.b); lane l of the quarter's
    // first wave sums partials l, l + 64, ... of it, then the wave sums its
    // lanes: the same order whatever the block shape, so a system's result
    // does not depend on the shape its batch runs in.
    int2 cr = make_int2(desc.y, desc.z);
    Batch<HV, batch_entries<BLOCK>()> cur;
    first_batch<WAVES>(cr.x, cr.y, rp, nbr_j, hv, cur);
    Staged<WAVES> sv;
    stage_load<WAVES>(mode, cr.x, cr.y, v, m_in, minv, n_all, binfo, eta, rhs, type, ff, sv, L);
    double pa = 0.0, pb = 0.0;
    const bool series_wave = threadIdx.x % QUARTER < 64;  // the first wave of its quarter
    if (series_wave) {
      const int series = (threadIdx.x / QUARTER) * 2;
      const double* pbase = series < NDOT ? part + ((k - 1) & 1) * NDOT * nd + series * nd : part_bb;
      pa = partial_sum<NP>(pbase + sd.dir0, sd.ndir);
      pb = partial_sum<NP>(pbase + nd + sd.dir0, sd.ndir);
    }
    __shared__ double red[8];
    // The first chunk: its sweep, then this launch's step from the reduced
    // partials, then its tail.  Straight-line code, so the first gathers wait
    // only for the first batch, not for the row operands and partials loaded
    // after it.
    chunk_spmv<WAVES, HV, VT, AT, IT>(cr.x, cr.y, rp, nbr_j, hv, vin, coff, cur, L.sums);
    stage_store<WAVES>(cr.x, cr.y, sv, L);
    if (series_wave) {
      pa = xr_wave_sum(pa);
      pb = xr_wave_sum(pb);
      if ((threadIdx.x & 63) == 0) {
        red[(threadIdx.x / QUARTER) * 2] = pa;
        red[(threadIdx.x / QUARTER) * 2 + 1] = pb;
      }
    }
    __syncthreads();
    {
      double res[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) res[q] = red[q];
      bool conv[2];
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        double target = tol2 * res[6 + r];
        if (MIXED && j > 0) target = fmax(target, INNER_DROP2 * c.rr0[r]);
        conv[r] = res[4 + r] <= target;
      }
      if ((conv[0] && conv[1]) || c.it >= max_iter || (MIXED && j >= INNER_MAX)) {
        if (lead) {
          int state;
          if (conv[0] && conv[1] && (!MIXED || j == 0)) state = CG_DONE;
          else if (c.it >= max_iter || c.restarts > MAX_RESTARTS) state = CG_STALLED;
          else state = CG_REFRESH;
          settle(&cg[sys], cnt, c, state, k);
        }
        return;
      }
      st.fresh = j == 0;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const double g = res[r], d = res[2 + r];
        double a = 0.0, b = 0.0;
        if (!conv[r]) {
          if (st.fresh || c.alp[r] == 0.0 || c.gam[r] == 0.0) {
            a = d != 0.0 ? g / d : 0.0;
          } else {
            b = g / c.gam[r];
            const double den = d - b * g / c.alp[r];
            a = den != 0.0 ? g / den : 0.0;
          }
        }
        st.alpha[r] = a;
        st.beta[r] = b;
      }
      if (lead) {
        CgCtl s = c;
        s.done = CG_RUNNING;
        if (st.fresh) {
          s.k0 = k;
          s.rr0[0] = res[4];
          s.rr0[1] = res[5];
        }
        s.it = c.it + 1;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          // a converged right-hand side keeps its last step (alpha = 0 stops it)
          s.gam[r] = conv[r] ? c.gam[r] : res[r];
          s.alp[r] = conv[r] ? c.alp[r] : st.alpha[r];
        }
        cg[sys].ctl[k & 1] = s;
      }
    }
    if (threadIdx.x < 128) chunk_tail<MIXED>(mode, cr.x, cr.y, st, v, m_out, m32_out, L, dots);
    // further chunks of a block (systems with more chunks than the block cap)
    for (int ch = c0 + 1; ch < c1; ++ch) {
      __syncthreads();
      cr = chunks[sd.upd0 + ch];
      rp = wave_row_ptrs<WAVES>(cr.x, cr.y, row_ptr);
      first_batch<WAVES>(cr.x, cr.y, rp, nbr_j, hv, cur);
      stage_load<WAVES>(mode, cr.x, cr.y, v, m_in, minv, n_all, binfo, eta, rhs, type, ff, sv, L);
      chunk_spmv<WAVES, HV, VT, AT, IT>(cr.x, cr.y, rp, nbr_j, hv, vin, coff, cur, L.sums);
      stage_store<WAVES>(cr.x, cr.y, sv, L);
      __syncthreads();
      if (threadIdx.x < 128) chunk_tail<MIXED>(mode, cr.x, cr.y, st, v, m_out, m32_out, L, dots);
    }
  } else {
    // refresh in FP64: r = b - A x and u = M^-1 r, then w = A u and m = M^-1 w.
    // A start that follows a residual launch first tests that true residual
    // (its r.r and b.b partials, reduced in the order an iteration launch
    // uses): a system that already meets the tolerance settles here, one
    // launch before the iteration launch would, with the same decision.
    if (mode == MODE_START && c.kr >= 0) {
      __shared__ double chk[4];  // r.r and b.b of both right-hand sides
      const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
      if (w < 4) {
        const double* src = w < 2 ? part + ((k - 1) & 1) * NDOT * nd + (4 + w) * nd : part_bb + (w - 2) * nd;
        double a = partial_sum<NP>(src + sd.dir0, sd.ndir);
        a = xr_wave_sum(a);
        if (l == 0) chk[w] = a;
      }
      __syncthreads();
      if (chk[0] <= tol2 * chk[2] && chk[1] <= tol2 * chk[3]) {
        if (lead) settle(&cg[sys], cnt, c, CG_DONE, k);
        return;
      }
    }
    if (lead) {
      CgCtl s = c;
      if (mode == MODE_RESIDUAL) s.restarts = c.restarts + 1;
      cg[sys].ctl[k & 1] = s;
    }
    const double2* vin = mode == MODE_RESIDUAL ? v.X : v.U;
    for (int ch = c0; ch < c1; ++ch) {
      const int2 cr = ch == c0 ? make_int2(desc.y, desc.z) : chunks[sd.upd0 + ch];
      if (ch > c0) rp = wave_row_ptrs<WAVES>(cr.x, cr.y, row_ptr);
      Batch<double, batch_entries<BLOCK>()> cur;
      first_batch<WAVES>(cr.x, cr.y, rp, nbr_j, hval, cur);
      Staged<WAVES> sv;
      stage_load<WAVES>(mode, cr.x, cr.y, v, m_in, minv, n_all, binfo, eta, rhs, type, ff, sv, L);
      chunk_spmv<WAVES, double, double2, double, IT>(cr.x, cr.y, rp, nbr_j, hval, vin, coff, cur, L.sums);
      stage_store<WAVES>(cr.x, cr.y, sv, L);
      __syncthreads();
      if (threadIdx.x < 128) chunk_tail<MIXED>(mode, cr.x, cr.y, st, v, m_out, m32_out, L, dots);
      if (ch + 1 < c1) __syncthreads();
    }
  }
  if (threadIdx.x < 128) {
    // wave h holds right-hand side h of every row: one wave sum per series,
    // lane 0 stores series 2q + h (after a residual launch only r.r, and b.b)
    const int h = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 4; ++q) dots[q] = xr_wave_sum(dots[q]);
    if ((threadIdx.x & 63) == 0) {
      double* out = part + (k & 1) * NDOT * nd;
      if (mode == MODE_RESIDUAL) {
        out[(4 + h) * nd + b] = dots[2];
      } else {
#pragma unroll
        for (int q = 0; q < 3; ++q) out[(2 * q + h) * nd + b] = dots[q];
      }
      if (mode != MODE_ITER) part_bb[h * nd + b] = dots[3];
    }
  }
}

// Per system: q = s + mu t with mu = (Q - sum s) / sum t; the (s, t) solution
// is kept in `keep` (when given) as the next call's starting guess.
__global__ void __launch_bounds__(1024) k_qeq_charges(Systems S, const double2* __restrict__ X,
                                                      double* __restrict__ q, double* __restrict__ mu_out,
                                                      double2* __restrict__ keep, int neutral) {
  const SysDev& sd = S.sys[blockIdx.x];
  __shared__ double red[2][16];
  // two independent partial sums per thread and sum, so the loads are in flight together
  double ss = 0.0, st = 0.0, ss2 = 0.0, st2 = 0.0;
  const int end = sd.atom0 + sd.n;
  int i = sd.atom0 + threadIdx.x;
  for (; i + 1024 < end; i += 2048) {
    const double2 a = X[i], b = X[i + 1024];
    ss += a.x;
    st += a.y;
    ss2 += b.x;
    st2 += b.y;
  }
  for (; i < end; i += 1024) {
    ss += X[i].x;
    st += X[i].y;
  }
  ss += ss2;
  st += st2;
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
                        const void* slab, int slab_stride, int* nbr_j, int* nbr_s, unsigned short* col16,
                        double bond_cut, double* hval, float* hval32, unsigned char* bflag, int* bond_count,
                        unsigned char* join, int2* hb_cand, int* hb_count, hipStream_t s) {
  const int blocks = (n + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;
  XR_LAUNCH(XT_EDGE, k_edge_prep, dim3(blocks), dim3(256), 0, s, n, x, type, S, ff, row_ptr,
            static_cast<const unsigned long long*>(slab), slab_stride, nbr_j, nbr_s, col16, bond_cut, hval, hval32,
            bflag, bond_count, join, hb_cand, hb_count);
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

// Chunks and iteration blocks per system: ceil(n / CHUNK) chunks (CHUNK_SMALL
// below CG_SMALL_ATOMS atoms), spread over at most CG_MAX_BLOCKS blocks
// (CG_MAX_BLOCKS_SMALL) with every block owning at least one chunk
// (XREAC_CG_MAX_BLOCKS overrides both caps for tuning; it must be the same for
// every call that is compared bitwise, since a system's blocks depend only on
// its size and this cap).
bool xr_cg_small(int n) { return n < CG_SMALL_ATOMS; }

void xr_cg_blocks(int n, int* ndir, int* nupd, int* chunk) {
  static const int cap_env = [] {
    const char* v = std::getenv("XREAC_CG_MAX_BLOCKS");
    const int c = v ? std::atoi(v) : 0;
    return c > 0 ? c : 0;
  }();
  const bool small = xr_cg_small(n);
  const int rows = small ? CHUNK_SMALL : CHUNK;
  const int cap = cap_env ? cap_env : small ? CG_MAX_BLOCKS_SMALL : CG_MAX_BLOCKS;
  *chunk = rows;
  const int chunks = n > 0 ? (n + rows - 1) / rows : 1;
  const int per = (chunks + cap - 1) / cap;
  *nupd = chunks;
  *ndir = (chunks + per - 1) / per;
}

size_t xr_cg_partials(int nd, int nu) {
  (void)nu;
  return (size_t)(2 * NDOT * nd + 2 * nd) + 64;
}

size_t xr_cg_block_doubles(int n) { return (size_t)n * BLOCK_MAX; }

size_t xr_cg_state_bytes(int ns) { return sizeof(CgSys) * (size_t)ns + sizeof(CgCounters) + 64; }

// One launch of iteration k over `blocks` iteration blocks of one shape,
// starting at block `base`; columns from the 2-byte copy when there is one.
template <bool MIXED, int BLOCK>
void cg_launch(const CgArgs& a, int k, double* part, double* part_bb, int nd, double tol2, CgSys* cg,
               CgCounters* cnt, int blocks, int base, hipStream_t s) {
  if (a.col16)
    XR_LAUNCH(XT_CG_ITERATION, (k_cg_iteration<MIXED, BLOCK, unsigned short>), dim3(blocks), dim3(BLOCK), 0, s, a.S,
              k, a.v, a.row_ptr, a.col16, a.hval, a.hval32, a.eta, a.minv, a.binfo, a.chunks, a.blocks, a.type, a.ff,
              a.rhs, part, part_bb, nd, tol2, a.max_iter, cg, cnt, a.n, base);
  else
    XR_LAUNCH(XT_CG_ITERATION, (k_cg_iteration<MIXED, BLOCK, int>), dim3(blocks), dim3(BLOCK), 0, s, a.S, k, a.v,
              a.row_ptr, a.nbr_j, a.hval, a.hval32, a.eta, a.minv, a.binfo, a.chunks, a.blocks, a.type, a.ff, a.rhs,
              part, part_bb, nd, tol2, a.max_iter, cg, cnt, a.n, base);
}

hipError_t xr_cg_solve(const CgArgs& a, hipStream_t s, CgOutcome* out) {
  const int nu = a.nu, nd = a.nd, ns = a.S.ns;
  double* part = a.partials;
  double* part_bb = part + 2 * NDOT * nd;
  CgCounters* cnt = static_cast<CgCounters*>(a.state);
  CgSys* cg = xr_cg_systems(a.state);
  const double tol2 = a.tol * a.tol;
  const bool mixed = a.hval32 != nullptr;
  // counters start at zero (cleared with the control block)
  XR_LAUNCH(XT_CG_INIT, k_cg_init, dim3(nu), dim3(64), 0, s, a.S, a.type, a.ff, a.rhs, a.guess, a.v, a.row_ptr,
            a.nbr_j, a.hval, a.join, a.minv, a.binfo, a.chunks, a.blocks, a.eta, cg, cnt, a.abort_flag, a.n);
  XR_LAUNCH_CHECK();
  int k = 0;
  // Every system settles within max_iter iterations plus three launches per
  // refresh (the settling launch, the true residual and the start).
  const long long k_limit = (long long)a.max_iter + 3LL * (MAX_RESTARTS + 3) + a.first_check + a.check_every;
  out->aborted = false;
  out->exhausted = false;
  // iteration blocks [0, nd_small) belong to small systems, the rest to large ones
  const int nd_small = a.nd_small, nd_large = nd - a.nd_small;
  const bool narrow = nd_small >= CG_NARROW_MIN_BLOCKS;
  auto enqueue = [&](int iters) {
    for (int c = 0; c < iters && k < k_limit; ++c) {
      if (mixed) {
        if (nd_small > 0 && narrow)
          cg_launch<true, CG_BLOCK_NARROW>(a, k, part, part_bb, nd, tol2, cg, cnt, nd_small, 0, s);
        else if (nd_small > 0)
          cg_launch<true, CG_BLOCK_SMALL>(a, k, part, part_bb, nd, tol2, cg, cnt, nd_small, 0, s);
        if (nd_large > 0) cg_launch<true, CG_BLOCK>(a, k, part, part_bb, nd, tol2, cg, cnt, nd_large, nd_small, s);
      } else {
        if (nd_small > 0 && narrow)
          cg_launch<false, CG_BLOCK_NARROW>(a, k, part, part_bb, nd, tol2, cg, cnt, nd_small, 0, s);
        else if (nd_small > 0)
          cg_launch<false, CG_BLOCK_SMALL>(a, k, part, part_bb, nd, tol2, cg, cnt, nd_small, 0, s);
        if (nd_large > 0) cg_launch<false, CG_BLOCK>(a, k, part, part_bb, nd, tol2, cg, cnt, nd_large, nd_small, s);
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
  auto charges = [&]() {
    hipLaunchKernelGGL(k_qeq_charges, dim3(ns), dim3(1024), 0, s, a.S, a.v.X, a.q, a.mu, a.keep, a.neutral ? 1 : 0);
    return hipGetLastError();
  };
  // A few launches go out before the caller's side-stream work, so the device
  // is busy while the host issues it, and the rest of the first chunk after.
  const int lead = a.after_start ? std::min(a.fixed_launches > 0 ? a.fixed_launches : a.first_check,
                                            CG_LEAD_LAUNCHES)
                                 : 0;
  hipError_t e = enqueue(lead);
  if (e == hipSuccess && a.after_start) e = a.after_start(a.after_start_arg);
  if (e != hipSuccess) return e;
  if (a.fixed_launches > 0) {
    // A recorded solve: a fixed number of launches and no read-back.  Systems
    // that settle earlier sit out the rest; the caller reads every system's
    // final state from the control block afterwards.
    e = enqueue(a.fixed_launches - lead);
    if (e == hipSuccess) e = charges();
    out->sweeps = k;
    out->tail_done = false;
    return e;
  }
  e = issue(0, a.first_check - lead);
  if (e != hipSuccess) return e;
  if (a.tail) {
    e = charges();
    if (e == hipSuccess) e = a.tail(a.tail_arg);
    if (e != hipSuccess) return e;
  }
  const int k_tail = k;
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
  out->tail_done = a.tail && k == k_tail && !out->aborted && !out->exhausted;
  if (!out->tail_done) return charges();
  return hipSuccess;
}

