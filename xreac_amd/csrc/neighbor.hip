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

// Visit every candidate image (j, T) of atom i in the surrounding bins; calls
// fn(j, packed_shift, r2) for those passing the exact distance test.  Bins
// with consecutive z index are contiguous in `binned`, so for each (x, y)
// offset the z-range is walked as one or a few contiguous segments (split only
// where the periodic image shift changes); lanes stride over each segment
// (lane = t, stride = nthreads).
template <class Fn>
__device__ inline void for_each_neighbor(int i, const double* __restrict__ x, const SysDev& sd,
                                         const int* __restrict__ home, const int* __restrict__ bin,
                                         const int* __restrict__ bin_start, const BinRec* __restrict__ recs,
                                         double rc2, int t, int nthreads, unsigned& err, Fn fn) {
  const CellDev& cell = sd.cell;
  const BinGeom& g = sd.g;
  const int id = bin[i] - sd.bin0;
  const int bz = id % g.nb[2], by = (id / g.nb[2]) % g.nb[1], bx = id / (g.nb[2] * g.nb[1]);
  const int hi0 = home[3 * i], hi1 = home[3 * i + 1], hi2 = home[3 * i + 2];
  const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
  // Position of atom i inside its bin (bin units, clamped to [0, 1]), the
  // same arithmetic as the binning pass.
  double a[3];
  {
    double f[3];
    frac_of(cell, x, i, f);
    const int bb[3] = {bx, by, bz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double u = g.pbc[k] ? (f[k] - floor(f[k])) / g.w[k] : (f[k] - g.lo[k]) / g.w[k];
      a[k] = fmin(1.0, fmax(0.0, u - bb[k]));
    }
  }
  // Lower bound on the distance from atom i to the slab of bins at offset o
  // along plane normal k (a small slack absorbs rounding of the bin
  // assignment).  A bin whose three bounds give sum d_k^2 > kappa rc^2 holds
  // no neighbor.
  const double lim2 = g.kappa * rc2;
  auto gap = [&](int k, int o) {
    const double d = o > 0 ? o - a[k] : (o < 0 ? a[k] - o - 1 : 0.0);
    return fmax(0.0, d * g.thick[k] - 1e-8);
  };
  for (int o0 = -g.reach[0]; o0 <= g.reach[0]; ++o0) {
    const double d0 = gap(0, o0);
    if (d0 * d0 > lim2) continue;
    int c0 = bx + o0, t0 = 0;
    if (g.pbc[0]) {
      t0 = floordiv(c0, g.nb[0]);
      c0 -= t0 * g.nb[0];
    } else if (c0 < 0 || c0 >= g.nb[0]) {
      continue;
    }
    for (int o1 = -g.reach[1]; o1 <= g.reach[1]; ++o1) {
      const double d1 = gap(1, o1);
      const double rem2 = lim2 - d0 * d0 - d1 * d1;
      if (rem2 < 0) continue;
      int c1 = by + o1, t1 = 0;
      if (g.pbc[1]) {
        t1 = floordiv(c1, g.nb[1]);
        c1 -= t1 * g.nb[1];
      } else if (c1 < 0 || c1 >= g.nb[1]) {
        continue;
      }
      // z offsets o with gap(2, o)^2 <= rem2
      const double rz = fmin((sqrt(rem2) + 1e-8) / g.thick[2], 1e6);
      const int olo = -(int)floor(rz + 1 - a[2]), ohi = (int)floor(rz + a[2]);
      int zlo = bz + (olo > -g.reach[2] ? olo : -g.reach[2]);
      int zhi = bz + (ohi < g.reach[2] ? ohi : g.reach[2]);
      if (!g.pbc[2]) {
        zlo = zlo < 0 ? 0 : zlo;
        zhi = zhi >= g.nb[2] ? g.nb[2] - 1 : zhi;
      }
      const int col = sd.bin0 + (c0 * g.nb[1] + c1) * g.nb[2];
      const int sx0 = g.pbc[0] ? t0 + hi0 : 0;
      const int sy0 = g.pbc[1] ? t1 + hi1 : 0;
      int zz = zlo;
      while (zz <= zhi) {
        // one segment: consecutive z bins sharing the same image shift t2
        const int t2 = g.pbc[2] ? floordiv(zz, g.nb[2]) : 0;
        const int cz0 = zz - t2 * g.nb[2];
        int last = zhi;
        if (g.pbc[2]) {
          const int seg_end = (t2 + 1) * g.nb[2] - 1;
          last = last < seg_end ? last : seg_end;
        }
        const int cz1 = last - t2 * g.nb[2];
        const int s = bin_start[col + cz0], e = bin_start[col + cz1 + 1];
        const int sz0 = g.pbc[2] ? t2 + hi2 : 0;
        // Every lane takes part in every round (hit = false past the end),
        // so fn may use wave-wide ballots.
        for (int base = s; base < e; base += nthreads) {
          const int p = base + t;
          bool hit = false;
          int j = 0, code = 0;
          double r2 = 0.0;
          if (p < e) {
            const BinRec rec = recs[p];
            j = rec.atom;
            int hj0, hj1, hj2;
            xr_unpack_shift(rec.home, hj0, hj1, hj2);
            const int sx = g.pbc[0] ? sx0 - hj0 : 0;
            const int sy = g.pbc[1] ? sy0 - hj1 : 0;
            const int sz = g.pbc[2] ? sz0 - hj2 : 0;
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
        zz = last + 1;
      }
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
  for_each_neighbor(i, x, S.sys[sys], home, bin, bin_start, binned, rc2, lane, 64, err,
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

// Sort the first `count` keys of s (lane-major, K keys per lane) ascending
// with a bitonic network held in registers: partners closer than K sit in the
// same lane, farther ones are exchanged with __shfl_xor.  One wave.
template <int K>
__device__ inline void wave_sort(unsigned long long* s, int count) {
  constexpr int N = 64 * K;
  const int lane = threadIdx.x & 63;
  unsigned long long v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int idx = lane * K + k;
    v[k] = idx < count ? s[idx] : ~0ull;
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
          const unsigned long long o = __shfl_xor(v[k], ls, 64);
          const bool keep_min = lower == up;
          v[k] = keep_min ? (v[k] < o ? v[k] : o) : (v[k] > o ? v[k] : o);
        }
      } else {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int q = k ^ stride;
          if (q > k) {
            const bool up = ((lane * K + k) & size) == 0;
            const unsigned long long a = v[k], b = v[q];
            const bool swap = (a > b) == up;
            v[k] = swap ? b : a;
            v[q] = swap ? a : b;
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) s[lane * K + k] = v[k];
}

__global__ void __launch_bounds__(64) k_nbr_fill(int n, const double* __restrict__ x, Systems S,
                                                 const int* __restrict__ home, const int* __restrict__ bin,
                                                 const int* __restrict__ bin_start, const BinRec* __restrict__ binned,
                                                 double rc2, const int* __restrict__ row_ptr, int* __restrict__ nbr_j,
                                                 int* __restrict__ nbr_s) {
  __shared__ unsigned long long keys[ROW_CAP];
  const int i = blockIdx.x;
  if (i >= n) return;
  const int lane = threadIdx.x & 63;
  const int begin = row_ptr[i], count = row_ptr[i + 1] - begin;
  if (count > ROW_CAP) {
    // Reported by the count pass.  The row still gets in-range entries (the
    // atom itself, zero shift) so later kernels index valid memory; the
    // system's results are discarded.
    for (int t2 = lane; t2 < count; t2 += 64) {
      nbr_j[begin + t2] = i;
      nbr_s[begin + t2] = xr_pack_shift(0, 0, 0);
    }
    return;
  }
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  unsigned err = 0;
  int cursor = 0;  // identical in every lane
  for_each_neighbor(i, x, S.sys[S.asys[i]], home, bin, bin_start, binned, rc2, lane, 64, err,
                    [&](bool hit, int j, int code, double) {
                      const unsigned long long m = __ballot(hit);
                      if (hit) {
                        const int p = cursor + __popcll(m & below);
                        if (p < ROW_CAP) keys[p] = ((unsigned long long)(unsigned)j << 32) | (unsigned)code;
                      }
                      cursor += __popcll(m);
                    });
  __syncthreads();
  if (count <= 128) {
    wave_sort<2>(keys, count);
  } else if (count <= 256) {
    wave_sort<4>(keys, count);
  } else if (count <= 512) {
    wave_sort<8>(keys, count);
  } else {
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
  for (int t2 = lane; t2 < count; t2 += 64) {
    const unsigned long long key = keys[t2];
    nbr_j[begin + t2] = (int)(key >> 32);
    nbr_s[begin + t2] = (int)(key & 0xffffffffu);
  }
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

hipError_t xr_neighbor_fill(int n, const double* x, const Systems& S, const int* home, const int* bin,
                            const int* bin_start, const void* binned, double cutoff, const int* row_ptr, int* nbr_j,
                            int* nbr_s, hipStream_t s) {
  XR_LAUNCH(XT_NBR_FILL, k_nbr_fill, dim3(n), dim3(64), 0, s, n, x, S, home, bin, bin_start,
            static_cast<const BinRec*>(binned), cutoff * cutoff, row_ptr, nbr_j, nbr_s);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_pack_supplied(int n, const Systems& S, const int* row_ptr, const int* in_j, const int* in_s,
                            int* nbr_j, int* nbr_s, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_supplied, dim3((n + 3) / 4), dim3(256), 0, s, n, S, row_ptr, in_j, in_s, nbr_j, nbr_s);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}
