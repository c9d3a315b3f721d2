// Shared definitions for the gfx950 ReaxFF kernels (wave64, FP64).
//
// Layout conventions (device memory):
//   positions      double[n*3]   (x0 y0 z0 x1 ...), caller-owned
//   types          int32[n]      force-field type index, caller-owned
//   neighbor CSR   row_ptr int32[n+1], nbr_j int32[m], nbr_s int32[m]
//                  rows in canonical (j, Sx, Sy, Sz) order; nbr_s packs the
//                  image shift as three 10-bit fields offset by 512.
//   bond CSR       bond_ptr int32[n+1] over the active bonds (raw BO >= cutoff)
//                  of each row, in canonical order; per-bond arrays are indexed
//                  by bond slot b.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define XR_WAVE 64

// ---------------------------------------------------------------- error bits
enum XrErr : unsigned {
  XR_ERR_COINCIDENT = 1u << 0,
  XR_ERR_SHIFT_RANGE = 1u << 1,
  XR_ERR_ROW_TOO_LONG = 1u << 2,
  XR_ERR_BOC3 = 1u << 3,
  XR_ERR_BOC4 = 1u << 4,
  XR_ERR_BOC5 = 1u << 5,
  XR_ERR_COLLINEAR = 1u << 6,
  XR_ERR_REVERSE = 1u << 7,
  XR_ERR_NONFINITE = 1u << 8,
  XR_ERR_CAPACITY = 1u << 9,  // a buffer sized from the previous call was too small
  XR_ERR_NEIGHBORS = 1u << 10,  // a supplied neighbor entry points outside its system or its shift is too large
  XR_ERR_POSITIONS = 1u << 11,  // non-finite coordinates along a nonperiodic direction
};

// ------------------------------------------------------- parameter layouts
enum AtomParam {
  AP_RS, AP_RPI, AP_RPP, AP_VAL, AP_VALE, AP_VALBOC, AP_VALVAL, AP_MASS, AP_CHI, AP_ETA,
  AP_GAMMA, AP_PLP2, AP_POVUN2, AP_POVUN5, AP_PVAL3, AP_PVAL5, AP_HBOND, AP_CARBON,
  AP_MASS_C12, AP_MASS_O16, AP_COUNT
};
enum PairParam {
  PP_BO1, PP_BO2, PP_BO3, PP_BO4, PP_BO5, PP_BO6, PP_RS, PP_RP, PP_RPP, PP_OVC, PP_V13,
  PP_BOC3, PP_BOC4, PP_BOC5, PP_DES, PP_DEP, PP_DEPP, PP_BE1, PP_BE2, PP_OVUN1, PP_D,
  PP_ALPHA, PP_GAMMAW, PP_RVDW, PP_RCORE, PP_ECORE, PP_ACORE, PP_GAMMA, PP_COUNT
};
enum { ANG_W = 7, TOR_W = 5, HB_W = 4, NGEN = 39 };

// Energy components, in the reference's COMPONENTS order.
enum Comp {
  C_BOND, C_ATOM, C_LONE, C_MOLECULE, C_ANGLE, C_PENALTY, C_COA, C_HBOND, C_TORSION, C_CONJ,
  C_VDW, C_COULOMB, C_EFIELD, C_QEQ, NCOMP
};

struct FFDev {
  int nt;                      // number of force-field atom types
  int vdw_type;                // 1 shielded, 2 inner wall, 3 both
  const double* atom;          // [nt][AP_COUNT]
  const double* pair;          // [nt*nt][PP_COUNT]
  const double* gen;           // [39]
  const int* ang_off;          // [nt^3 + 1] rows for (t_i, t_center, t_k)
  const double* ang;           // [rows][7] (theta00 v1 v2 coa1 v7 pen1 v4), |v1|>0.001
  const double* tor;           // [nt^4][5] keyed (t_i, t_j, t_k, t_l)
  const unsigned char* tor_on; // [nt^4]
  const unsigned char* tor_mid;    // [nt^2] (t_j, t_k): some torsion row exists
  const unsigned char* tor_left;   // [nt^3] (t_i, t_j, t_k): some row exists for some t_l
  const unsigned char* tor_right;  // [nt^3] (t_j, t_k, t_l): some row exists for some t_i
  const double* hb;            // [nt^3][4] keyed (t_donor, t_H, t_acceptor)
  const unsigned char* hb_on;  // [nt^3]
  const double* vdw_b;         // [nt*nt] (1/gamma_w)^p_vdw, the shielding offset of each pair
  const double* vdw_c;         // [nt*nt][4] 1/r_vdw, -alpha/(2 r_vdw), 1/r_core, -a_core/r_core
};

struct CellDev {
  double L[9];     // rows are lattice vectors (identity when nonperiodic)
  double inv[9];   // inverse of L
  int pbc[3];
  int periodic;
};

// Cell-list grid of one system.
struct BinGeom {
  int nb[3];     // bins per direction
  int reach[3];  // bins searched on each side
  double lo[3];  // grid origin in fractional (periodic) or scaled units
  double w[3];   // bin width in the same units
  int pbc[3];
  double thick[3];  // Cartesian distance between the faces of a bin along each plane normal
  double kappa;     // bound on the largest eigenvalue of N N^T, N = unit plane normals (1: orthogonal)
};

// One system of a batch.  Atoms of system s occupy [atom0, atom0 + n) of the
// packed per-atom arrays; its bins occupy [bin0, bin0 + nb0*nb1*nb2) of the
// global bin arrays; its CG iteration blocks are [dir0, dir0 + ndir) (small
// systems' blocks first) and its chunks [upd0, upd0 + nupd).
struct SysDev {
  CellDev cell;
  BinGeom g;
  int atom0, n;
  int bin0;
  int bin_budget;  // bins reserved at bin0; nonperiodic directions are fitted into it on the device
  int dir0, ndir;
  int upd0, nupd;
  int chunk;  // nominal rows per CG chunk (by system size)
  double total_charge;
};

// Batch view shared by the kernels.
struct Systems {
  const SysDev* sys;  // [ns]
  const int* asys;    // [N] system of each atom
  unsigned* err;      // [ns] input-error bits per system
  int ns;
};

__device__ inline const CellDev& xr_cell(const Systems& S, int i) { return S.sys[S.asys[i]].cell; }
__device__ inline void xr_flag(const Systems& S, int i, unsigned bits) { atomicOr(&S.err[S.asys[i]], bits); }

// Per-edge flag byte written by the edge pass.
enum EdgeFlag : unsigned char {
  EF_BOND = 1,  // raw bond order at or above the cutoff: an active bond
};

// ------------------------------------------------------------ shift packing
__host__ __device__ inline int xr_pack_shift(int sx, int sy, int sz) {
  return ((sx + 512) << 20) | ((sy + 512) << 10) | (sz + 512);
}
__host__ __device__ inline void xr_unpack_shift(int code, int& sx, int& sy, int& sz) {
  sx = ((code >> 20) & 1023) - 512;
  sy = ((code >> 10) & 1023) - 512;
  sz = (code & 1023) - 512;
}
__host__ __device__ inline int xr_neg_shift(int code) {
  int sx, sy, sz;
  xr_unpack_shift(code, sx, sy, sz);
  return xr_pack_shift(-sx, -sy, -sz);
}

// S @ L with the reference's rounding: fma(S2, L2k, fma(S1, L1k, S0*L0k)).
__device__ inline void xr_offset(const CellDev& c, int code, double off[3]) {
  if (!c.periodic) {
    off[0] = off[1] = off[2] = 0.0;
    return;
  }
  int sx, sy, sz;
  xr_unpack_shift(code, sx, sy, sz);
  const double s0 = (double)sx, s1 = (double)sy, s2 = (double)sz;
#pragma unroll
  for (int k = 0; k < 3; ++k) off[k] = fma(s2, c.L[6 + k], fma(s1, c.L[3 + k], s0 * c.L[k]));
}

// Edge vector (x_j - x_i) + off and its length, rounded as the reference does:
// |v|^2 = (v0*v0 + v1*v1) + v2*v2, no contraction (file compiled with
// -ffp-contract=off).
__device__ inline double xr_edge(const double* __restrict__ x, const CellDev& c, int i, int j, int code,
                                 double v[3]) {
  double off[3];
  xr_offset(c, code, off);
#pragma unroll
  for (int k = 0; k < 3; ++k) v[k] = (x[3 * j + k] - x[3 * i + k]) + off[k];
  return sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
}

// ------------------------------------------------------------ wave helpers
__device__ inline double xr_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, XR_WAVE);
  return v;  // valid in lane 0
}
// Sum over aligned groups of G lanes (G a power of two <= 64), valid in the
// group's first lane; the same tree as xr_wave_sum restricted to the group.
template <int G>
__device__ inline double xr_group_sum(double v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_down(v, off, G);
  return v;
}
__device__ inline int xr_wave_sum_int(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, XR_WAVE);
  return v;
}
__device__ inline unsigned xr_lane() { return __lane_id(); }

// Orders a wave's LDS writes before its later LDS reads by other lanes of the
// same wave (no block barrier: waves of a block run independent loops).
__device__ inline void xr_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Block reduction of NV doubles into partials[k * nblocks + blockIdx.x].
// Fixed tree order -> deterministic.  Block size must be a multiple of 64.
template <int NV>
__device__ inline void xr_block_partials(double (&vals)[NV], double* __restrict__ partials, int nblocks) {
  __shared__ double red[NV][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = xr_wave_sum(vals[k]);
    if (lane == 0) red[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[threadIdx.x][w];
    partials[threadIdx.x * nblocks + blockIdx.x] = s;
  }
  __syncthreads();
}

__device__ inline double xr_sq(double a) { return a * a; }

// Host-side launch check helper.
#define XR_LAUNCH_CHECK() \
  do {                    \
    hipError_t e__ = hipGetLastError(); \
    if (e__ != hipSuccess) return e__;  \
  } while (0)

// ------------------------------------------------- uncorrected bond orders
// sigma' = (1+c) exp(p_bo1 (r/r_s)^p_bo2), pi', pipi' similarly, each gated on
// both atoms having a positive radius (src/xreac/energy.py:130-139).
struct BondRaw {
  double sig, pi, pp;     // values
  double dsig, dpi, dpp;  // derivatives with respect to r
};

__device__ inline void xr_bond_raw(const FFDev& ff, int ti, int tj, double r, BondRaw& o) {
  const double* ai = ff.atom + ti * AP_COUNT;
  const double* aj = ff.atom + tj * AP_COUNT;
  const double* p = ff.pair + (ti * ff.nt + tj) * PP_COUNT;
  const double c = 0.01 * ff.gen[29];
  o.sig = o.pi = o.pp = o.dsig = o.dpi = o.dpp = 0.0;
  if (ai[AP_RS] > 0 && aj[AP_RS] > 0) {
    const double q = pow(r / p[PP_RS], p[PP_BO2]);
    o.sig = (1 + c) * exp(p[PP_BO1] * q);
    o.dsig = o.sig * p[PP_BO1] * p[PP_BO2] * q / r;
  }
  if (ai[AP_RPI] > 0 && aj[AP_RPI] > 0) {
    const double q = pow(r / p[PP_RP], p[PP_BO4]);
    o.pi = exp(p[PP_BO3] * q);
    o.dpi = o.pi * p[PP_BO3] * p[PP_BO4] * q / r;
  }
  if (ai[AP_RPP] > 0 && aj[AP_RPP] > 0) {
    const double q = pow(r / p[PP_RPP], p[PP_BO6]);
    o.pp = exp(p[PP_BO5] * q);
    o.dpp = o.pp * p[PP_BO5] * p[PP_BO6] * q / r;
  }
}

// Taper (energy.py:90-91): 1 + u^4(-35 + u(84 + u(-70 + 20u))), u = r/rc, 0 at r >= rc.
__device__ inline void xr_taper(double r, double rc, double& t, double& dt) {
  if (r < rc) {
    const double inv_rc = 1.0 / rc;  // loop-invariant in the edge loops: hoisted
    const double u = r * inv_rc;
    const double u2 = u * u;
    const double u4 = u2 * u2;
    const double inner = -35 + u * (84 + u * (-70 + 20 * u));
    t = 1 + u4 * inner;
    // d/du [u^4 inner] = 4u^3 inner + u^4 (84 - 140u + 60u^2)
    const double du = 4 * (u2 * u) * inner + u4 * (84 + u * (-140 + 60 * u));
    dt = du * inv_rc;
  } else {
    t = 0.0;
    dt = 0.0;
  }
}

// Shielded Coulomb kernel shield = taper / (r^3 + gamma_ij)^(1/3) (energy.py:92),
// from a taper value t and its derivative dt already at hand.
__device__ inline void xr_shield_tapered(double r, double t, double dt, double gamma, double& s, double& ds) {
  const double r2 = r * r;
  const double base = r2 * r + gamma;
  const double inv = rcbrt(base);  // base^(-1/3)
  s = t * inv;
  // d/dr base^(1/3) = r^2 base^(1/3) / base, and 1 / base = inv^3
  ds = (dt - t * (r2 * ((inv * inv) * inv))) * inv;
}

__device__ inline void xr_shield(double r, double rc, double gamma, double& s, double& ds) {
  double t, dt;
  xr_taper(r, rc, t, dt);
  xr_shield_tapered(r, t, dt, gamma, s, ds);
}
