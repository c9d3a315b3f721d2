// Exclusive prefix sum of int32 counts (row sizes -> CSR row pointers).
//
// Up to ONE_MAX counts: one launch of one block that walks the input in tiles
// with a running carry, and can clamp the pointers to a capacity and take the
// largest count on the way.  Larger inputs: three launches (per-block totals,
// one-block scan of the totals, per-block scan with the block offset added),
// then the clamp and the maximum as launches of their own.  out has n + 1
// entries; out[n] = total.
#include "common.h"
#include "launch.h"

namespace {
constexpr int ONE_BLOCK = 1024;
constexpr int ONE_ITEMS = 16;
constexpr int ONE_TILE = ONE_BLOCK * ONE_ITEMS;
constexpr int ONE_MAX = 4 * ONE_TILE;  // 65,536 counts
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ inline int block_exclusive_scan(int v, int* tmp, int& total) {
  // Inclusive scan within each wave, then across waves.
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) tmp[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      int t = tmp[w];
      tmp[w] = acc;
      acc += t;
    }
    tmp[16] = acc;
  }
  __syncthreads();
  int excl = x - v + tmp[wave];
  total = tmp[16];
  __syncthreads();
  return excl;
}

__global__ void k_tile_sums(const int* __restrict__ in, int n, int* __restrict__ sums) {
  __shared__ int tmp[17];
  const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) s += (base + k < n) ? in[base + k] : 0;
  int total;
  block_exclusive_scan(s, tmp, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void k_scan_sums(int* __restrict__ sums, int nt) {
  __shared__ int tmp[17];
  int carry = 0;
  for (int base = 0; base < nt; base += blockDim.x) {
    int idx = base + threadIdx.x;
    int v = idx < nt ? sums[idx] : 0;
    int total;
    int e = block_exclusive_scan(v, tmp, total);
    if (idx < nt) sums[idx] = carry + e;
    carry += total;
  }
  if (threadIdx.x == 0) sums[nt] = carry;
}

__global__ void k_tile_scan(const int* __restrict__ in, int n, const int* __restrict__ sums, int* __restrict__ out) {
  __shared__ int tmp[17];
  const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int e = block_exclusive_scan(s, tmp, total) + sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) out[base + k] = e;
    e += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = sums[gridDim.x];
}
// Largest count in a block: the wave tree, then the waves in order (thread 0 holds it).
__device__ inline int block_max(int m) {
  __shared__ int red[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_down(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = red[w] > m ? red[w] : m;
  return m;
}

// One block, n <= ONE_MAX: each thread scans 16 consecutive counts of a tile,
// the block scans the thread sums, and a carry runs from tile to tile; the
// next tile's counts are loaded before the current tile is scanned.
// Pointers above cap are written as cap; the unclamped total goes to *total
// and raises XR_ERR_CAPACITY when it exceeds cap.  The largest count goes to
// *max_out.  (total, err, max_out may be null.)
__global__ void __launch_bounds__(ONE_BLOCK) k_scan_one(const int* __restrict__ in, int n, int* __restrict__ out,
                                                        int cap, int* __restrict__ total_out,
                                                        unsigned* __restrict__ err, int* __restrict__ max_out) {
  __shared__ int tmp[17];
  int carry = 0, mx = 0;
  int nv[ONE_ITEMS];
  // A thread's 16 counts are four 16-byte loads when all lie inside a 16-byte
  // aligned input, else guarded single loads.
  auto load = [&](int t0) {
    const int base = t0 + threadIdx.x * ONE_ITEMS;
    if (base + ONE_ITEMS <= n && (reinterpret_cast<size_t>(in) & 15) == 0) {
      const int4* v4 = reinterpret_cast<const int4*>(in + base);
#pragma unroll
      for (int q = 0; q < ONE_ITEMS / 4; ++q) {
        const int4 w = v4[q];
        nv[4 * q] = w.x;
        nv[4 * q + 1] = w.y;
        nv[4 * q + 2] = w.z;
        nv[4 * q + 3] = w.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < ONE_ITEMS; ++k) nv[k] = base + k < n ? in[base + k] : 0;
    }
  };
  load(0);
  for (int t0 = 0; t0 < n; t0 += ONE_TILE) {
    const int base = t0 + threadIdx.x * ONE_ITEMS;
    int v[ONE_ITEMS];
#pragma unroll
    for (int k = 0; k < ONE_ITEMS; ++k) v[k] = nv[k];
    if (t0 + ONE_TILE < n) load(t0 + ONE_TILE);
    int s = 0;
#pragma unroll
    for (int k = 0; k < ONE_ITEMS; ++k) {
      s += v[k];
      mx = v[k] > mx ? v[k] : mx;
    }
    int total;
    int e = block_exclusive_scan(s, tmp, total) + carry;
    int o[ONE_ITEMS];
#pragma unroll
    for (int k = 0; k < ONE_ITEMS; ++k) {
      o[k] = e < cap ? e : cap;
      e += v[k];
    }
    if (base + ONE_ITEMS <= n && (reinterpret_cast<size_t>(out + base) & 15) == 0) {
      int4* o4 = reinterpret_cast<int4*>(out + base);
#pragma unroll
      for (int q = 0; q < ONE_ITEMS / 4; ++q) o4[q] = make_int4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
      for (int k = 0; k < ONE_ITEMS; ++k)
        if (base + k < n) out[base + k] = o[k];
    }
    carry += total;
  }
  if (threadIdx.x == 0) {
    out[n] = carry < cap ? carry : cap;
    if (total_out) *total_out = carry;
    if (err && carry > cap) atomicOr(err, (unsigned)XR_ERR_CAPACITY);
  }
  if (max_out) {
    mx = block_max(mx);
    if (threadIdx.x == 0) *max_out = mx;
  }
}

__global__ void k_clamp_ptr(int* __restrict__ ptr, int n, int cap, int* __restrict__ total,
                            unsigned* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const int v = ptr[i];
  if (i == n) {
    *total = v;
    if (v > cap) atomicOr(err, (unsigned)XR_ERR_CAPACITY);
  }
  if (v > cap) ptr[i] = cap;
}

__global__ void __launch_bounds__(1024) k_max_count(int n, const int* __restrict__ in, int* __restrict__ out) {
  int m = 0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) m = in[t] > m ? in[t] : m;
  m = block_max(m);
  if (threadIdx.x == 0) *out = m;
}
}  // namespace

size_t xr_scan_scratch_ints(int n) { return (size_t)((n + SCAN_TILE - 1) / SCAN_TILE) + 2; }

hipError_t xr_exclusive_scan(const int* in, int n, int* out, int* scratch, hipStream_t s) {
  if (n <= ONE_MAX) {
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(ONE_BLOCK), 0, s, in, n, out, INT32_MAX, nullptr, nullptr,
                       nullptr);
    XR_LAUNCH_CHECK();
    return hipSuccess;
  }
  const int tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(SCAN_BLOCK), 0, s, in, n, scratch);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, scratch, tiles);
  hipLaunchKernelGGL(k_tile_scan, dim3(tiles), dim3(SCAN_BLOCK), 0, s, in, n, scratch, out);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_clamp_ptr(int* ptr, int n, int cap, int* total, unsigned* err, hipStream_t s) {
  hipLaunchKernelGGL(k_clamp_ptr, dim3((n + 1 + 255) / 256), dim3(256), 0, s, ptr, n, cap, total, err);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}

hipError_t xr_scan_clamped(const int* in, int n, int* out, int* scratch, int cap, int* total, unsigned* err,
                           int* max_out, hipStream_t s) {
  if (n <= ONE_MAX) {
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(ONE_BLOCK), 0, s, in, n, out, cap, total, err, max_out);
    XR_LAUNCH_CHECK();
    return hipSuccess;
  }
  hipError_t e = xr_exclusive_scan(in, n, out, scratch, s);
  if (e == hipSuccess) e = xr_clamp_ptr(out, n, cap, total, err, s);
  if (e == hipSuccess && max_out) {
    hipLaunchKernelGGL(k_max_count, dim3(1), dim3(1024), 0, s, n, in, max_out);
    XR_LAUNCH_CHECK();
  }
  return e;
}
