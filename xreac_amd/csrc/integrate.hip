// On-device integrators for packed batches: velocity Verlet (NVE) and FIRE.
//
// One block per system; the per-system reductions (power, norms, largest
// step, kinetic energy) are taken in a fixed order, so a system's trajectory
// does not depend on the batch it runs in.
//
// FIRE restates the reference's native relaxation loop
// (src/xreac/calculator.py:213-246, Bitzek et al., PRL 97, 170201 (2006)):
// unit fictitious masses, dt0 = 0.02, dt_max = 0.2, alpha0 = 0.1, growth 1.1
// after more than five positive-power steps, alpha *= 0.99, halving on
// negative power, displacement capped at 0.1 A per atom and iteration.
//
// Velocity Verlet, and the extrapolated QEq guesses of its steps, have no
// counterpart in the reference (it has no dynamics).
// Units: A, fs, amu, kcal/mol; F/m in kcal/mol/A/amu times 4.184e-4 is A/fs^2.
#include "integrate.h"

namespace {

constexpr double ACCEL = 4.184e-4;  // (kcal/mol/A/amu) -> A/fs^2
constexpr int IBLK = 256;

template <class Op>
__device__ inline double block_reduce(double v, Op op, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = red[0];
  for (int k = 1; k < (int)(blockDim.x >> 6); ++k) r = op(r, red[k]);
  __syncthreads();
  return r;
}

struct Add {
  __device__ double operator()(double a, double b) const { return a + b; }
};
struct Max {
  __device__ double operator()(double a, double b) const { return fmax(a, b); }
};

// stage 0: v += dt/2 a; x += dt v.   stage 1: v += dt/2 a; kinetic energy.
__global__ void __launch_bounds__(IBLK) k_verlet(const int* __restrict__ offsets, const int* __restrict__ type,
                                                 const double* __restrict__ atom, const double* __restrict__ dt_fs,
                                                 double* __restrict__ x, double* __restrict__ v,
                                                 const double* __restrict__ F, int stage,
                                                 double* __restrict__ kinetic) {
  __shared__ double red[IBLK / 64];
  const int s = blockIdx.x;
  const int a0 = offsets[s], a1 = offsets[s + 1];
  const double dt = dt_fs[s];
  const double half = 0.5 * dt * ACCEL;
  double ke = 0.0;
  for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    const double m = atom[type[i] * AP_COUNT + AP_MASS];
    const double c = half / m;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double vk = v[3 * i + k] + c * F[3 * i + k];
      v[3 * i + k] = vk;
      if (stage == 0) x[3 * i + k] += dt * vk;
      ke += m * vk * vk;
    }
  }
  if (stage == 1 && kinetic) {
    ke = block_reduce(ke, Add{}, red);
    if (threadIdx.x == 0) kinetic[s] = 0.5 * ke / ACCEL;
  }
}

// Starting guess of the QEq solve of the next MD step: the polynomial of
// degree k through the system's last k + 1 solutions (equal time steps),
// evaluated one step ahead.  With h0 the current solution and h1..h3 the
// previous ones (newest first), the guess is sum_q (-1)^q C(k+1, q+1) h_q:
//   k = 1: 2 h0 - h1,  k = 2: 3 (h0 - h1) + h2,  k = 3: 4 (h0 + h2) - (6 h1 + h3),
// and k = 0 keeps h0.  The history then shifts by one (h0 becomes h1).
__global__ void __launch_bounds__(IBLK) k_qeq_extrapolate(const int* __restrict__ offsets,
                                                          const int* __restrict__ order, double2* __restrict__ state,
                                                          double2* __restrict__ hist, int n) {
  const int s = blockIdx.x;
  const int a0 = offsets[s], a1 = offsets[s + 1];
  const int k = min(max(order[s], 0), 3);
  double2* h1p = hist;
  double2* h2p = hist + n;
  double2* h3p = hist + 2 * (size_t)n;
  for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    const double2 h0 = state[i], h1 = h1p[i], h2 = h2p[i], h3 = h3p[i];
    double2 g = h0;
    if (k == 1) {
      g.x = 2.0 * h0.x - h1.x;
      g.y = 2.0 * h0.y - h1.y;
    } else if (k == 2) {
      g.x = 3.0 * (h0.x - h1.x) + h2.x;
      g.y = 3.0 * (h0.y - h1.y) + h2.y;
    } else if (k == 3) {
      g.x = 4.0 * (h0.x + h2.x) - (6.0 * h1.x + h3.x);
      g.y = 4.0 * (h0.y + h2.y) - (6.0 * h1.y + h3.y);
    }
    h3p[i] = h2;
    h2p[i] = h1;
    h1p[i] = h0;
    state[i] = g;
  }
}

// One FIRE iteration (calculator.py:225-245).  state[s] = (dt, alpha,
// positive steps, iteration); status[s]: 0 running, 1 converged, 2 limit.
__global__ void __launch_bounds__(IBLK) k_fire(const int* __restrict__ offsets, double* __restrict__ x,
                                               double* __restrict__ v, const double* __restrict__ F,
                                               double* __restrict__ state, int* __restrict__ status, double ftol,
                                               int max_iterations) {
  __shared__ double red[IBLK / 64];
  __shared__ double sh[4];
  const int s = blockIdx.x;
  if (status[s] != 0) return;
  const int a0 = offsets[s], a1 = offsets[s + 1];
  double fbig = 0.0, power = 0.0, vv = 0.0, ff = 0.0;
  for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double f = F[3 * i + k], vk = v[3 * i + k];
      fbig = fbig > fabs(f) ? fbig : fabs(f);
      power += vk * f;
      vv += vk * vk;
      ff += f * f;
    }
  }
  fbig = block_reduce(fbig, Max{}, red);
  power = block_reduce(power, Add{}, red);
  vv = block_reduce(vv, Add{}, red);
  ff = block_reduce(ff, Add{}, red);
  if (threadIdx.x == 0) {
    double dt = state[4 * s], alpha = state[4 * s + 1];
    int positive = (int)state[4 * s + 2], it = (int)state[4 * s + 3];
    int st = 0;
    double mix = 0.0, keep = 1.0;  // v <- keep * v + mix * F before the dt F kick
    if (fbig <= ftol) {
      st = 1;
    } else if (it >= max_iterations) {
      st = 2;
    } else if (power > 0) {
      keep = 1 - alpha;
      mix = alpha * sqrt(vv) / sqrt(ff);
      positive += 1;
      if (positive > 5) {
        dt = fmin(dt * 1.1, 0.2);
        alpha *= 0.99;
      }
    } else {
      keep = 0.0;
      positive = 0;
      alpha = 0.1;
      if (it > 0) dt *= 0.5;
    }
    if (st) status[s] = st;
    else {
      state[4 * s] = dt;
      state[4 * s + 1] = alpha;
      state[4 * s + 2] = positive;
      state[4 * s + 3] = it + 1;
    }
    sh[0] = st;
    sh[1] = keep;
    sh[2] = mix;
    sh[3] = dt;
  }
  __syncthreads();
  if (sh[0] != 0.0) return;
  const double keep = sh[1], mix = sh[2], dt = sh[3];
  // new velocities; the largest displacement |dt v_i| (row norm as numpy takes it)
  double largest = 0.0;
  for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    double d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double f = F[3 * i + k];
      double vk = keep == 0.0 ? 0.0 : keep * v[3 * i + k] + mix * f;
      vk += dt * f;
      v[3 * i + k] = vk;
      d[k] = dt * vk;
    }
    largest = fmax(largest, sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]));
  }
  largest = block_reduce(largest, Max{}, red);
  const bool cap = largest > 0.1;
  const double scale = cap ? 0.1 / largest : 1.0;
  for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double d = dt * v[3 * i + k];
      if (cap) d *= scale;
      x[3 * i + k] += d;
    }
  }
}

}  // namespace

hipError_t xr_verlet(const FFDev& ff, int ns, const int* offsets, const int* types, const double* dt_fs, double* x,
                     double* v, const double* F, int stage, double* kinetic, hipStream_t s) {
  hipLaunchKernelGGL(k_verlet, dim3(ns), dim3(IBLK), 0, s, offsets, types, ff.atom, dt_fs, x, v, F, stage, kinetic);
  return hipGetLastError();
}

hipError_t xr_qeq_extrapolate(int ns, int n, const int* offsets, const int* order, double* state, double* hist,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_qeq_extrapolate, dim3(ns), dim3(IBLK), 0, s, offsets, order, reinterpret_cast<double2*>(state),
                     reinterpret_cast<double2*>(hist), n);
  return hipGetLastError();
}

hipError_t xr_fire(int ns, const int* offsets, double* x, double* v, const double* F, double* state, int* status,
                   double ftol, int max_iterations, hipStream_t s) {
  hipLaunchKernelGGL(k_fire, dim3(ns), dim3(IBLK), 0, s, offsets, x, v, F, state, status, ftol, max_iterations);
  return hipGetLastError();
}
