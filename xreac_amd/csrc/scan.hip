// Exclusive prefix sum of int32 counts (row sizes -> CSR row pointers).
//
// Three launches: per-block totals, one-block scan of the totals, per-block
// scan with the block offset added.  out has n + 1 entries; out[n] = total.
#include "common.h"
#include "launch.h"

namespace {
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
}  // namespace

size_t xr_scan_scratch_ints(int n) { return (size_t)((n + SCAN_TILE - 1) / SCAN_TILE) + 2; }

hipError_t xr_exclusive_scan(const int* in, int n, int* out, int* scratch, hipStream_t s) {
  const int tiles = n > 0 ? (n + SCAN_TILE - 1) / SCAN_TILE : 1;
  hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(SCAN_BLOCK), 0, s, in, n, scratch);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, scratch, tiles);
  hipLaunchKernelGGL(k_tile_scan, dim3(tiles), dim3(SCAN_BLOCK), 0, s, in, n, scratch, out);
  XR_LAUNCH_CHECK();
  return hipSuccess;
}
