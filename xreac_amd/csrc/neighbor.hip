// Cell-list neighbor build producing the canonical directed list.
//
// Membership is the reference's exact test (src/xreac/neighbors.py:41-46):
// (i, j, S) is an edge iff ((x_j - x_i) + S@L)^2 <= rc^2 and (i, j, S) is not
// (i, i, 0).  Binning only proposes candidates; the test itself is evaluated on
// the caller's (possibly unwrapped) coordinates with the reference rounding.
//
// Bins live in fractional coordinates for periodic directions (each bin at
// least rc/2 thick measured along the face normal) and in scaled coordinates
// for nonperiodic ones.  In a batch every system has its own grid, placed at
// its own offset in the global bin arrays; atoms only ever meet atoms of their
// own system.  Rows are sorted in LDS by the 64-bit key
// (j << 32) | packed(S), which is the reference's (i, j, Sx, Sy, Sz)
// lexicographic order within row i (neighbors.py:65-98).
#include "launch.h"
#include "profile.h"

namespace {

constexpr int ROW_CAP = 2048;  // max neighbors per atom held in LDS for sorting

__device__ inline void frac_of(const CellDev& c, const double* x, int i, double f[3]) {
  const double a = x[3 * i], b = x[3 * i + 1], d = x[3 * i + 2];
#pragma unroll
  for (int k = 0; k < 3; ++k) f[k] = a * c.inv[k] + b * c.inv[3 + k] + d * c.inv[6 + k];
}

__device__ inline int floordiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Bin-ordered copy of an atom: exact coordinates, index and its home image
// (three 10-bit fields offset by 512), 32 bytes.
struct BinRec {
  double x, y, z;
  int atom;
  int home;
};

// One block per system: fractional-coordinate bounds of its atoms.
__global__ void __launch_bounds__(256) k_coord_bounds(Systems S, const double* __restrict__ x,
                                                      double* __restrict__ out) {
  const SysDev& sd = S.sys[blockIdx.x];
  if (sd.cell.pbc[0] && sd.cell.pbc[1] && sd.cell.pbc[2]) return;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = sd.atom0 + threadIdx.x; i < sd.atom0 + sd.n; i += blockDim.x) {
    double f[3];
    frac_of(sd.cell, x, i, f);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = fmin(lo[k], f[k]);
      hi[k] = fmax(hi[k], f[k]);
    }
  }
  __shared__ double red[6][256];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = lo[k];
    red[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + st]);
        red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + st]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// Bin grid along the nonperiodic directions of each system, from its
// coordinate bounds (fractional units of its cell), one thread per system:
// bins at least rc/2 thick (at most 128 per direction), widened by steps of
// 1.25 until the system's bins fit the budget reserved for it on the host
// (the periodic directions' counts are fixed there).  Non-finite bounds flag
// the system and leave one bin per direction, so later kernels stay in bounds.
__global__ void k_bin_grid(SysDev* __restrict__ sys, int ns, const double* __restrict__ bounds, double rcb,
                           unsigned* __restrict__ err) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ns) return;
  SysDev& sd = sys[q];
  BinGeom& g = sd.g;
  if (g.pbc[0] && g.pbc[1] && g.pbc[2]) return;
  const double* b = bounds + 6 * q;
  double height[3], span[3], w[3];
  long long nb[3];
  int periodic = 1;
  bool finite = true;
  for (int k = 0; k < 3; ++k) {
    const double c0 = sd.cell.inv[k], c1 = sd.cell.inv[3 + k], c2 = sd.cell.inv[6 + k];
    height[k] = 1.0 / sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    if (g.pbc[k]) {
      periodic *= g.nb[k];
      continue;
    }
    span[k] = b[3 + k] - b[k];
    finite = finite && isfinite(span[k]) && isfinite(b[k]);
  }
  if (!finite) {
    atomicOr(&err[q], (unsigned)XR_ERR_POSITIONS);
    for (int k = 0; k < 3; ++k) {
      if (g.pbc[k]) continue;
      g.nb[k] = 1;
      g.w[k] = 1.0;
      g.lo[k] = 0.0;
      g.reach[k] = 1;
      g.thick[k] = height[k];
    }
    return;
  }
  const long long budget = sd.bin_budget / periodic > 0 ? sd.bin_budget / periodic : 1;
  double scale = 1.0;
  for (int round = 0; round < 200; ++round) {
    long long total = 1;
    for (int k = 0; k < 3; ++k) {
      if (g.pbc[k]) continue;
      w[k] = 0.5 * rcb / height[k] * scale;
      nb[k] = (long long)floor(span[k] / w[k]) + 1;
      if (nb[k] > 128) {
        w[k] = span[k] / 127.0;
        nb[k] = 128;
      }
      total *= nb[k];
    }
    if (total <= budget) break;
    scale *= 1.25;
  }
  for (int k = 0; k < 3; ++k) {
    if (g.pbc[k]) continue;
    g.nb[k] = (int)nb[k];
    g.w[k] = w[k];
    g.lo[k] = b[k];
    g.reach[k] = (int)floor(rcb / (w[k] * height[k])) + 1;
    g.thick[k] = w[k] * height[k];
  }
}

__global__ void k_bin_atoms(int n, const double* __restrict__ x, Systems S, int* __restrict__ home,
                            int* __restrict__ bin, int* __restrict__ bin_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SysDev& sd = S.sys[S.asys[i]];
  const BinGeom& g = sd.g;
  double f[3];
  frac_of(sd.cell, x, i, f);
  int b[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double u;
    if (g.pbc[k]) {
      const double h = floor(f[k]);
      home[3 * i + k] = (int)h;
      u = (f[k] - h) / g.w[k];
    } else {
      home[3 * i + k] = 0;
      u = (f[k] - g.lo[k]) / g.w[k];
    }
    int bk = (int)floor(u);
    b[k] = bk < 0 ? 0 : (bk >= g.nb[k] ? g.nb[k] - 1 : bk);
  }
  const int id = sd.bin0 + (b[0] * g.nb[1] + b[1]) * g.nb[2] + b[2];
  bin[i] = id;
  atomicAdd(&bin_count[id], 1);
}

__global__ void k_bin_scatter(int n, const double* __restrict__ x, const int* __restrict__ home,
                              const int* __restrict__ bin, const int* __restrict__ bin_start,
                              int* __restrict__ bin_fill, BinRec* __restrict__ recs, Systems S) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = bin[i];
  const int h0 = home[3 * i], h1 = home[3 * i + 1], h2 = home[3 * i + 2];
  if (h0 < -511 || h0 > 511 || h1 < -511 || h1 > 511 || h2 < -511 || h2 > 511) xr_flag(S, i, XR_ERR_SHIFT_RANGE);
  BinRec r;
  r.x = x[3 * i];
  r.y = x[3 * i + 1];
  r.z = x[3 * i + 2];
  r.atom = i;
  r.home = xr_pack_shift(h0, h1, h2);
  recs[bin_start[b] + atomicAdd(&bin_fill[b], 1)] = r;
}

// Visit every candidate image (j, T) of atom i in the bins around it and call
// fn(hit, j, packed_shift, r2) for each.  Bins that can hold a neighbor (the
// slab-distance bound below) are taken 64 at a time, one per lane: each lane
// reads its bin's record range, a wave prefix sum lays the ranges end to end,
// and the lanes stride over the joined list, so one round of 64 candidates
// spans several bins and one latency of range loads serves 64 bins.  Every
// lane calls fn in every round (hit = false past the end), so fn may use
// wave-wide ballots.
template <class Fn>
__device__ inline void for_each_neighbor(int i, const double* __restrict__ x, const SysDev& sd,
                                         const int* __restrict__ home, const int* __restrict__ bin,
                                         const int* __restrict__ bin_start, const BinRec* __restrict__ recs,
                                         double rc2, unsigned& err, Fn fn) {
  const int lane = threadIdx.x & 63;
  const CellDev& cell = sd.cell;
  const BinGeom& g = sd.g;
  const int id = bin[i] - sd.bin0;
  const int bb[3] = {id / (g.nb[2] * g.nb[1]), (id / g.nb[2]) % g.nb[1], id % g.nb[2]};
  const int hi[3] = {home[3 * i], home[3 * i + 1], home[3 * i + 2]};
  const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
  // Position of atom i inside its bin (bin units, clamped to [0, 1]), the
  // same arithmetic as the binning pass.
  double a[3];
  {
    double f[3];
    frac_of(cell, x, i, f);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double u = g.pbc[k] ? (f[k] - floor(f[k])) / g.w[k] : (f[k] - g.lo[k]) / g.w[k];
      a[k] = fmin(1.0, fmax(0.0, u - bb[k]));
    }
  }
  // Lower bound on the distance from atom i to the bins at offset o along
  // plane normal k (a small slack absorbs rounding of the bin assignment).  A
  // bin whose bounds give sum_k d_k^2 > kappa rc^2 holds no neighbor.
  const double lim2 = g.kappa * rc2;
  const int span1 = 2 * g.reach[1] + 1, span2 = 2 * g.reach[2] + 1;
  const int nbins = (2 * g.reach[0] + 1) * span1 * span2;
  for (int b0 = 0; b0 < nbins; b0 += 64) {
    // this lane's bin: candidate range [start, start + len) and image base
    const int q = b0 + lane;
    int start = 0, len = 0;
    int base[3] = {0, 0, 0};
    if (q < nbins) {
      const int o[3] = {q / (span1 * span2) - g.reach[0], (q / span2) % span1 - g.reach[1], q % span2 - g.reach[2]};
      double d2 = 0.0;
      int c[3];
      bool ok = true;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double d = o[k] > 0 ? o[k] - a[k] : (o[k] < 0 ? a[k] - o[k] - 1 : 0.0);
        const double gk = fmax(0.0, d * g.thick[k] - 1e-8);
        d2 += gk * gk;
        c[k] = bb[k] + o[k];
        if (g.pbc[k]) {
          const int t = floordiv(c[k], g.nb[k]);
          c[k] -= t * g.nb[k];
          base[k] = t + hi[k];
        } else if (c[k] < 0 || c[k] >= g.nb[k]) {
          ok = false;
        }
      }
      if (ok && d2 <= lim2) {
        const int cell_id = sd.bin0 + (c[0] * g.nb[1] + c[1]) * g.nb[2] + c[2];
        start = bin_start[cell_id];
        len = bin_start[cell_id + 1] - start;
      }
    }
    // inclusive prefix of the range lengths over the lanes
    int incl = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int v = __shfl_up(incl, off, 64);
      if (lane >= off) incl += v;
    }
    const int count = __shfl(incl, 63, 64);
    const int excl = incl - len;
    for (int r0 = 0; r0 < count; r0 += 64) {
      const int f = r0 + lane;
      // owner lane: the number of lanes whose ranges end at or before f
      int sl = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (__shfl(incl, sl + step - 1, 64) <= f) sl += step;
      sl = sl < 63 ? sl : 63;
      const int p = __shfl(start, sl, 64) + (f - __shfl(excl, sl, 64));
      const int s0 = __shfl(base[0], sl, 64), s1 = __shfl(base[1], sl, 64), s2 = __shfl(base[2], sl, 64);
      bool hit = false;
      int j = 0, code = 0;
      double r2 = 0.0;
      if (f < count) {
        const BinRec rec = recs[p];
        j = rec.atom;
        int hj0, hj1, hj2;
        xr_unpack_shift(rec.home, hj0, hj1, hj2);
        const int sx = g.pbc[0] ? s0 - hj0 : 0;
        const int sy = g.pbc[1] ? s1 - hj1 : 0;
        const int sz = g.pbc[2] ? s2 - hj2 : 0;
        if (sx < -511 || sx > 511 || sy < -511 || sy > 511 || sz < -511 || sz > 511) {
          err |= XR_ERR_SHIFT_RANGE;
        } else {
          code = xr_pack_shift(sx, sy, sz);
          double off[3];
          xr_offset(cell, code, off);
          const double v0 = (rec.x - xi0) + off[0];
          const double v1 = (rec.y - xi1) + off[1];
          const double v2 = (rec.z - xi2) + off[2];
          r2 = (v0 * v0 + v1 * v1) + v2 * v2;
          hit = r2 <= rc2 && !(j == i && sx == 0 && sy == 0 && sz == 0);
        }
      }
      fn(hit, j, code, r2);
    }
  }
}

__global__ void __launch_bounds__(256) k_nbr_count(int n, const double* __restrict__ x, Systems S,
                                                   const int* __restrict__ home, const int* __restrict__ bin,
                                                   const int* __restrict__ bin_start, const BinRec* __restrict__ binned,
                                                   double rc2, int row_cap, int* __restrict__ counts) {
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int sys = __builtin_amdgcn_readfirstlane(S.asys[i]);
  unsigned err = 0;
  int cnt = 0;
  for_each_neighbor(i, x, S.sys[sys], home, bin, bin_start, binned, rc2, err,
                    [&](bool hit, int j, int, double r2) {
                      if (hit) {
                        ++cnt;
                        if (j != i && r2 < 1e-12) err |= XR_ERR_COINCIDENT;
                      }
                    });
  cnt = xr_wave_sum_int(cnt);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) err |= __shfl_down(err, off, 64);
  if (lane == 0) {
    counts[i] = cnt;
    if (cnt > row_cap) err |= XR_ERR_ROW_TOO_LONG;
    if (err) atomicOr(&S.err[sys], err);
  }
}

// Sort `count` keys (lane-major, K keys per lane) ascending with a bitonic
// network held in registers: partners closer than K sit in the same lane,
// farther ones are exchanged with __shfl_xor.  One wave; load(idx) gives key
// idx < count, store(idx, key) takes the sorted key idx.
template <int K, class T, class Load, class Store>
__device__ inline void wave_sort(int count, Load load, Store store) {
  constexpr int N = 64 * K;
  const int lane = threadIdx.x & 63;
  T v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int idx = lane * K + k;
    v[k] = idx < count ? load(idx) : T(~T(0));
  }
#pragma unroll
  for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= K) {
        const int ls = stride / K;
        const bool lower = (lane & ls) == 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const bool up = ((lane * K + k) & size) == 0;
          const T o = __shfl_xor(v[k], ls, 64);
          const bool keep_min = lower == up;
          v[k] = keep_min ? (v[k] < o ? v[k] : o) : (v[k] > o ? v[k] : o);
        }
      } else {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int q = k ^ stride;
          if (q > k) {
            const bool up = ((lane * K + k) & size) == 0;
            const T a = v[k], b = v[q];
            const bool swap = (a > b) == up;
            v[k] = swap ? b : a;
            v[q] = swap ? a : b;
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int idx = lane * K + k;
    if (idx < count) store(idx, v[k]);
  }
}

// A row whose image shifts all lie in {-1, 0, 1} (any cell taller than the
// cutoff) sorts 32-bit keys (j - atom0) << 5 | (9 (sx + 1) + 3 (sy + 1) +
// (sz + 1)), ordered as the full (j << 32 | packed shifts) keys.
__device__ inline bool compact_shift(int code) {
  return (unsigned)(((code >> 20) & 1023) - 511) <= 2u && (unsigned)(((code >> 10) & 1023) - 511) <= 2u &&
         (unsigned)((code & 1023) - 511) <= 2u;
}

__device__ inline unsigned compact_key(unsigned long long key, int atom0) {
  const int code = (int)(key & 0xffffffffu);
  const int img = 9 * (((code >> 20) & 1023) - 511) + 3 * (((code >> 10) & 1023) - 511) + ((code & 1023) - 511);
  return ((unsigned)((int)(key >> 32) - atom0) << 5) | (unsigned)img;
}

__device__ inline unsigned long long full_key(unsigned c, int atom0) {
  const int img = (int)(c & 31u);
  const int j = (int)(c >> 5) + atom0;
  const int code = xr_pack_shift(img / 9 - 1, (img / 3) % 3 - 1, img % 3 - 1);
  return ((unsigned long long)(unsigned)j << 32) | (unsigned)code;
}

// One wave per row: the row's neighbors, sorted into canonical (j, image)
// order, go to the row's slab slots [i * CAP, i * CAP + count) as 64-bit keys
// (j << 32 | image code) and the count to counts[i]; a scan of the counts then
// gives the CSR row pointers, and the edge pass moves the keys into place.
// The same walk flags coincident atoms and over-long rows.  Rows are staged in
// LDS of CAP keys: the 512-key (4 KB) and 1024-key (8 KB) instances sort in
// registers and leave room for many waves per CU; a launch that meets a row
// longer than its CAP drops the row (count 0) and raises the capacity bit, and
// the evaluation is repeated with the instance the longest row needs.
// With a gate (one word per system), rows of systems whose word is 0 are left
// as they are: a kept skin list that is still valid.
template <int CAP>
__global__ void __launch_bounds__(64) k_nbr_fill(int n, const double* __restrict__ x, Systems S,
                                                 const int* __restrict__ home, const int* __restrict__ bin,
                                                 const int* __restrict__ bin_start, const BinRec* __restrict__ binned,
                                                 double rc2, int* __restrict__ counts,
                                                 unsigned long long* __restrict__ slab, unsigned* __restrict__ flags,
                                                 const unsigned* __restrict__ gate) {
  __shared__ unsigned long long keys[CAP];
  const int i = blockIdx.x;
  if (i >= n) return;
  const int lane = threadIdx.x & 63;
  const int sys = __builtin_amdgcn_readfirstlane(S.asys[i]);
  if (gate && gate[sys] == 0u) return;
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  unsigned err = 0;
  int cursor = 0;  // identical in every lane
  bool wide = false;  // a shift outside {-1, 0, 1}: full 64-bit sort keys
  for_each_neighbor(i, x, S.sys[sys], home, bin, bin_start, binned, rc2, err,
                    [&](bool hit, int j, int code, double r2) {
                      const unsigned long long m = __ballot(hit);
                      if (hit) {
                        const int p = cursor + __popcll(m & below);
                        if (p < CAP) keys[p] = ((unsigned long long)(unsigned)j << 32) | (unsigned)code;
                        if (j != i && r2 < 1e-12) err |= XR_ERR_COINCIDENT;
                        if (!compact_shift(code)) wide = true;
                      }
                      cursor += __popcll(m);
                    });
  const int atom0 = S.sys[sys].atom0;
  const bool compact = __ballot(wide) == 0ull && S.sys[sys].n <= (1 << 27);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) err |= __shfl_down(err, off, 64);
  const int count = cursor;
  if (count > CAP) {
    // Over ROW_CAP: the system's results are discarded; over CAP only: the
    // evaluation is repeated with the long-row instance.  The row is left
    // empty, so later kernels stay in bounds.
    if (lane == 0) {
      counts[i] = 0;
      if (count > ROW_CAP) err |= XR_ERR_ROW_TOO_LONG;
      else atomicOr(flags, (unsigned)XR_ERR_CAPACITY);
      if (err) atomicOr(&S.err[sys], err);
    }
    return;
  }
  if (lane == 0) {
    counts[i] = count;
    if (err) atomicOr(&S.err[sys], err);
  }
  __syncthreads();
  unsigned long long* out = slab + (size_t)i * CAP;
  // Register sorts store straight to the row's slab slots.
  auto ld64 = [&](int idx) { return keys[idx]; };
  auto st64 = [&](int idx, unsigned long long key) { out[idx] = key; };
  auto ld32 = [&](int idx) { return compact_key(keys[idx], atom0); };
  auto st32 = [&](int idx, unsigned key) { out[idx] = full_key(key, atom0); };
  if (count <= 512 || (CAP == 1024 && count <= 1024)) {
    if (compact) {
      if (count <= 128) wave_sort<2, unsigned>(count, ld32, st32);
      else if (count <= 256) wave_sort<4, unsigned>(count, ld32, st32);
      else if (count <= 512) wave_sort<8, unsigned>(count, ld32, st32);
      else if constexpr (CAP == 1024) wave_sort<16, unsigned>(count, ld32, st32);
    } else {
      if (count <= 128) wave_sort<2, unsigned long long>(count, ld64, st64);
      else if (count <= 256) wave_sort<4, unsigned long long>(count, ld64, st64);
      else if (count <= 512) wave_sort<8, unsigned long long>(count, ld64, st64);
      else if constexpr (CAP == 1024) wave_sort<16, unsigned long long>(count, ld64, st64);
    }
    return;
  }
  if constexpr (CAP > 1024) {
    // Long rows: bitonic network in LDS.
    int size = 1024;
    while (size < count) size <<= 1;
    for (int t2 = count + lane; t2 < size; t2 += 64) keys[t2] = ~0ull;
    __syncthreads();
    for (int k = 2; k <= size; k <<= 1) {
      for (int jj = k >> 1; jj > 0; jj >>= 1) {
        for (int t2 = lane; t2 < size; t2 += 64) {
          const int partner = t2 ^ jj;
          if (partner > t2) {
            const unsigned long long a = keys[t2], b = keys[partner];
            if ((a > b) == ((t2 & k) == 0)) {
              keys[t2] = b;
              keys[partner] = a;
            }
          }
        }
        __syncthreads();
      }
    }
  }
  __syncthreads();
  for (int t2 = lane; t2 < count; t2 += 64) out[t2] = keys[t2];
}

// ---------------------------------------------------------- skin reuse
// A list kept from an earlier call holds, per atom, every image within
// rc + 2 skin of the positions it was built at (xref), sorted canonically.
// While no atom of a system has moved more than skin since then (and no type
// has changed), every pair now within rc is in it, so filtering it by the
// current distances gives exactly the rows a fresh build gives.

// gate[s] = 1 when system s needs a new skin list: forced by the host, an atom
// moved by more than sqrt(lim2) (or to a non-finite position), or a type changed.
__global__ void __launch_bounds__(256) k_skin_check(Systems S, const double* __restrict__ x,
                                                    const int* __restrict__ types, const double* __restrict__ xref,
                                                    const int* __restrict__ tref, double lim2, int force,
                                                    unsigned* __restrict__ gate) {
  const SysDev& sd = S.sys[blockIdx.x];
  __shared__ int moved;
  if (threadIdx.x == 0) moved = 0;
  __syncthreads();
  bool m = false;
  if (!force) {
    for (int i = sd.atom0 + threadIdx.x; i < sd.atom0 + sd.n; i += blockDim.x) {
      const double d0 = x[3 * i] - xref[3 * i], d1 = x[3 * i + 1] - xref[3 * i + 1],
                   d2 = x[3 * i + 2] - xref[3 * i + 2];
      m = m || !((d0 * d0 + d1 * d1) + d2 * d2 <= lim2) || types[i] != tref[i];
    }
  }
  if (m) moved = 1;  // every writer stores the same value
  __syncthreads();
  if (threadIdx.x == 0) gate[blockIdx.x] = (force || moved) ? 1u : 0u;
}

// Positions and types of the systems whose skin list was just rebuilt.
__global__ void k_skin_commit(int n, Systems S, const double* __restrict__ x, const int* __restrict__ types,
                              double* __restrict__ xref, int* __restrict__ tref, const unsigned* __restrict__ gate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || gate[S.asys[i]] == 0u) return;
  xref[3 * i] = x[3 * i];
  xref[3 * i + 1] = x[3 * i + 1];
  xref[3 * i + 2] = x[3 * i + 2];
  tref[i] = types[i];
}

// One wave per atom: the kept skin row filtered to rc with the fill's own
// distance arithmetic, in its canonical order, to the row's slab slots and
// its length to counts; coincident pairs and over-long rows are flagged as
// the fill flags them.
__global__ void __launch_bounds__(256) k_skin_filter(int n, const double* __restrict__ x, Systems S,
                                                     const unsigned long long* __restrict__ skin, int skin_stride,
                                                     const int* __restrict__ skin_counts, double rc2,
                                                     int* __restrict__ counts, unsigned long long* __restrict__ slab,
                                                     int stride, unsigned* __restrict__ flags) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int sys = __builtin_amdgcn_readfirstlane(S.asys[i]);
  const SysDev& sd = S.sys[sys];
  const CellDev& cell = sd.cell;
  const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
  const int zero = xr_pack_shift(0, 0, 0);
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int have = min(max(skin_counts[i], 0), skin_stride);
  const unsigned long long* row = skin + (size_t)i * skin_stride;
  unsigned long long* out = slab + (size_t)i * stride;
  unsigned err = 0;
  int cursor = 0;  // identical in every lane
  for (int b = 0; b < have; b += 64) {
    const int e = b + lane;
    bool hit = false;
    unsigned long long key = 0;
    if (e < have) {
      key = row[e];
      const int j = (int)(key >> 32), code = (int)(key & 0xffffffffu);
      if ((unsigned)(j - sd.atom0) < (unsigned)sd.n) {
        double off[3];
        xr_offset(cell, code, off);
        const double v0 = (x[3 * j] - xi0) + off[0];
        const double v1 = (x[3 * j + 1] - xi1) + off[1];
        const double v2 = (x[3 * j + 2] - xi2) + off[2];
        const double r2 = (v0 * v0 + v1 * v1) + v2 * v2;
        hit = r2 <= rc2 && !(j == i && code == zero);
        if (hit && j != i && r2 < 1e-12) err |= XR_ERR_COINCIDENT;
      } else {
        err |= XR_ERR_NEIGHBORS;  // not a row of this system: the kept list is not valid
      }
    }
    const unsigned long long m = __ballot(hit);
    if (hit) {
      const int p = cursor + __popcll(m & below);
      if (p < stride) out[p] = key;
    }
    cursor += __popcll(m);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) err |= __shfl_down(err, off, 64);
  if (lane == 0) {
    if (cursor > stride) {
      counts[i] = 0;  // the row is dropped and the evaluation repeated (or the system failed)
      if (cursor > ROW_CAP) err |= XR_ERR_ROW_TOO_LONG;
      else atomicOr(flags, (unsigned)XR_ERR_CAPACITY);
    } else {
      counts[i] = cursor;
    }
    if (err) atomicOr(&S.err[sys], err);
  }
}

// Supplied list: clamp every row pointer to [0, cap], so every later row
// walk stays inside the edge buffers whatever the caller passed.  A first
// pointer other than 0, a negative pointer or a pointer below its
// predecessor flags the row's system (such a row then walks nothing).  The
// clamp is idempotent, so a thread that reads a neighbour's pointer before or
// after that neighbour's thread clamped it compares the same two values.
__global__ void k_clamp_supplied(int* __restrict__ ptr, int n, int cap, int* __restrict__ total,
                                 unsigned* __restrict__ err, Systems S) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const int v = ptr[i];
  const int c = v < 0 ? 0 : (v > cap ? cap : v);
  bool bad = v < 0 || (i == 0 && v != 0);
  if (i < n) {
    const int w = ptr[i + 1];
    const int cw = w < 0 ? 0 : (w > cap ? cap : w);
    bad |= cw < c;
  }
  if (i == n) {
    *total = v;
    if (v > cap) atomicOr(err, (unsigned)XR_ERR_CAPACITY);
  }
  if (bad && n > 0) xr_flag(S, i < n ? i : n - 1, XR_ERR_NEIGHBORS);
  if (c != v) ptr[i] = c;
}

// Supplied list: copy j and pack the shifts into the device codes, one wave
// per row.  An entry whose atom lies outside the row's system or whose shift
// does not fit 10 bits is replaced by the row atom with zero shift (so later
// kernels stay in bounds) and flags the system.
__global__ void __launch_bounds__(256) k_pack_supplied(int n, Systems S, const int* __restrict__ row_ptr,
                                                       const int* __restrict__ in_j,
                                                       const int* __restrict__ in_s, int* __restrict__ nbr_j,
                                                       int* __restrict__ nbr_s) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const SysDev& sd = S.sys[S.asys[i]];
  bool bad = false;
  for (int e = row_ptr[i] + lane; e < row_ptr[i + 1]; e += 64) {
    int j = in_j[e];
    const int sx = in_s[3 * e], sy = in_s[3 * e + 1], sz = in_s[3 * e + 2];
    const bool ok = j >= sd.atom0 && j < sd.atom0 + sd.n && sx >= -511 && sx <= 511 && sy >= -511 && sy <= 511 &&
                    sz >= -511 && sz <= 511;
    nbr_j[e] = ok ? j : i;
    nbr_s[e] = ok ? xr_pack_shift(sx, sy, sz) : xr_pack_shift(0, 0, 0);
    bad |= !ok;
  }
  if (__ballot(bad) && lane == 0) xr_flag(S, i, XR_ERR_NEIGHBORS);
}

}  // namespace

hipError_t xr_coord_bounds(const Systems& S, const double* x, double* bounds, hipStream_t s) {
  hipLaunchKernelGGL(k_coord_bounds, dim3(S.ns), dim3(256), 0, s, S, x, bounds);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_bin_grid(SysDev* sys, int ns, const double* bounds, double rcb, unsigned* err, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_grid, dim3((ns + 63) / 64), dim3(64), 0, s, sys, ns, bounds, rcb, err);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_bin_atoms(int n, const double* x, const Systems& S, int* home, int* bin, int* bin_count,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_bin_atoms, dim3((n + 255) / 256), dim3(256), 0, s, n, x, S, home, bin, bin_count);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

size_t xr_bin_record_bytes() { return sizeof(BinRec); }

hipError_t xr_bin_scatter(int n, const double* x, const Systems& S, const int* home, const int* bin,
                          const int* bin_start, int* bin_fill, void* recs, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_scatter, dim3((n + 255) / 256), dim3(256), 0, s, n, x, home, bin, bin_start, bin_fill,
                     static_cast<BinRec*>(recs), S);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_neighbor_count(int n, const double* x, const Systems& S, const int* home, const int* bin,
                             const int* bin_start, const void* binned, double cutoff, int row_cap, int* counts,
                             hipStream_t s) {
  // row_cap < 0: count only (coincidence and image checks), no row-length limit
  const int cap = row_cap < 0 ? 0x7fffffff : (row_cap < ROW_CAP ? row_cap : ROW_CAP);
  XR_LAUNCH(XT_NBR_COUNT, k_nbr_count, dim3((n + 3) / 4), dim3(256), 0, s, n, x, S, home, bin, bin_start,
            static_cast<const BinRec*>(binned), cutoff * cutoff, cap, counts);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

int xr_fill_stride(int max_row) { return max_row <= 512 ? 512 : (max_row <= 1024 ? 1024 : ROW_CAP); }

hipError_t xr_neighbor_fill(int n, const double* x, const Systems& S, const int* home, const int* bin,
                            const int* bin_start, const void* binned, double cutoff, int* counts, void* slab,
                            int max_row, unsigned* flags, hipStream_t s, const unsigned* gate) {
  unsigned long long* keys = static_cast<unsigned long long*>(slab);
  const int stride = xr_fill_stride(max_row);
  if (stride == 512) {
    XR_LAUNCH(XT_NBR_FILL, k_nbr_fill<512>, dim3(n), dim3(64), 0, s, n, x, S, home, bin, bin_start,
              static_cast<const BinRec*>(binned), cutoff * cutoff, counts, keys, flags, gate);
  } else if (stride == 1024) {
    XR_LAUNCH(XT_NBR_FILL, k_nbr_fill<1024>, dim3(n), dim3(64), 0, s, n, x, S, home, bin, bin_start,
              static_cast<const BinRec*>(binned), cutoff * cutoff, counts, keys, flags, gate);
  } else {
    XR_LAUNCH(XT_NBR_FILL, k_nbr_fill<ROW_CAP>, dim3(n), dim3(64), 0, s, n, x, S, home, bin, bin_start,
              static_cast<const BinRec*>(binned), cutoff * cutoff, counts, keys, flags, gate);
  }
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_skin_check(const Systems& S, const double* x, const int* types, const double* xref, const int* tref,
                         double allowance, bool force, unsigned* gate, hipStream_t s) {
  hipLaunchKernelGGL(k_skin_check, dim3(S.ns), dim3(256), 0, s, S, x, types, xref, tref, allowance * allowance,
                     force ? 1 : 0, gate);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_skin_commit(int n, const Systems& S, const double* x, const int* types, double* xref, int* tref,
                          const unsigned* gate, hipStream_t s) {
  hipLaunchKernelGGL(k_skin_commit, dim3((n + 255) / 256), dim3(256), 0, s, n, S, x, types, xref, tref, gate);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_skin_filter(int n, const double* x, const Systems& S, const void* skin, int skin_stride,
                          const int* skin_counts, double cutoff, int* counts, void* slab, int stride, unsigned* flags,
                          hipStream_t s) {
  XR_LAUNCH(XT_NBR_FILL, k_skin_filter, dim3((n + 3) / 4), dim3(256), 0, s, n, x, S,
            static_cast<const unsigned long long*>(skin), skin_stride, skin_counts, cutoff * cutoff, counts,
            static_cast<unsigned long long*>(slab), stride, flags);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_clamp_supplied(int* row_ptr, int n, int cap, int* total, unsigned* err, const Systems& S,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_clamp_supplied, dim3((n + 1 + 255) / 256), dim3(256), 0, s, row_ptr, n, cap, total, err, S);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_pack_supplied(int n, const Systems& S, const int* row_ptr, const int* in_j, const int* in_s,
                            int* nbr_j, int* nbr_s, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_supplied, dim3((n + 3) / 4), dim3(256), 0, s, n, S, row_ptr, in_j, in_s, nbr_j, nbr_s);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}
