// Device state of one evaluation: per-bond and per-atom arrays, forward values
// and their adjoints.  All arrays are carved from one workspace allocation.
#pragma once
#include "common.h"

// Energy accumulation channels; several channels feed one reported component.
enum Channel {
  CH_BOND, CH_TRIPLE, CH_LP, CH_C2, CH_ATOM, CH_ANGLE, CH_PEN, CH_COA, CH_HB, CH_TOR, CH_CONJ, CH_VDW, CH_COUL,
  CH_QEQ, NCH
};

struct EnergyAcc {
  double* e;  // [NCH][n] energy of each channel attributed to each atom (zeroed before each evaluation)
  int n;
};

// Store atom i's share of a channel (one writer per channel and atom).
__device__ inline void xr_energy(const EnergyAcc& a, int ch, int i, double v) { a.e[(size_t)ch * a.n + i] = v; }

struct Bonds {
  int cap;                 // allocated bond slots (launch size of per-bond kernels)
  const int* nbp;          // device: active directed bonds counted by the bond scan, unclamped: above cap
                           // when the bond buffers overflowed (the call is then repeated); use xr_bond_count
  const int* ptr;          // [n+1]
  const int* edge;         // [nb] canonical edge index
  const int* bi;           // [nb] row atom
  int* bj;                 // [nb] far atom
  int* code;               // [nb] packed image shift
  int* rev;                // [nb] reverse bond slot
  double* vec;             // [nb*3] edge vector x_j - x_i + S@L
  double* r;               // [nb]
  double* bop;             // [nb] uncorrected bo' = raw - cutoff
  double* pim;             // [nb] uncorrected pi'
  double* ppm;             // [nb] uncorrected pipi'
  double* dsig, *dpi, *dpp;   // d/dr of sigma', pi', pipi'
  double* f1, *f4, *f5;    // correction factors
  double* boc, *pic, *ppc; // corrected, before the 1e-10 threshold
  double* BO, *PI, *PP, *SIG; // thresholded outputs
  // adjoints
  double* A_BO, *A_PI, *A_PP, *A_SIG, *A_totfar;
  double* G;               // [nb*3] dE/d(edge vector)
  double* a_bop, *a_pim, *a_ppm, *c_i, *c_j;
};

// Bond slots a per-bond kernel may touch: the active bonds, at most the
// allocated slots (a grid of cap threads rounded up to whole blocks must not
// reach past the arrays when the count overflowed them).
__device__ inline int xr_bond_count(const Bonds& bd) { return min(*bd.nbp, bd.cap); }

struct Atoms {
  int n;
  Systems S;               // systems of the batch (cells, per-system error bits)
  const double* x;
  const int* type;
  double* totp;            // uncorrected total bond order
  double* total, *d, *db, *nlp, *vlpex, *dlp, *dlpt, *sbo, *sbo2, *prod, *ov2;
  int* bond_count;         // bonds with BO > 0.3
  double *A_total, *A_d, *A_db, *A_nlp, *A_dlpt, *A_dlp, *A_sbo2, *a_ov2, *a_totp;
  double* bomax;           // largest corrected bond order over the atom's bonds (torsion pruning)
  double* F_nb;            // [n*3] nonbonded forces
  double* F_hb;            // [n*3] hydrogen-bond acceptor-vector forces
  double* q;               // charges
};

// bondorder.hip
hipError_t xr_bond_forward(const Atoms& at, const Bonds& bd, FFDev ff, const int* nbr_j, const int* nbr_s,
                           EnergyAcc acc, hipStream_t s);
hipError_t xr_bond_backward(const Atoms& at, const Bonds& bd, FFDev ff, hipStream_t s);
// Final per-atom forces from the bond gradients and the nonbonded and H-bond forces.
hipError_t xr_force_assembly(const Atoms& at, const Bonds& bd, double* forces, hipStream_t s);
// manybody.hip
// hb_cand / hb_count: the edge pass's H-bond candidate lists (xr_edge_prep).
hipError_t xr_manybody(const Atoms& at, const Bonds& bd, FFDev ff, const int* row_ptr, const int2* hb_cand,
                       const int* hb_count, int lammps_angles, EnergyAcc acc, int* long_atoms, unsigned* long_count, hipStream_t s);
// pairs.hip
// van der Waals (positions only; second stream): forces written into F_hb, before the H-bond kernels add theirs.
hipError_t xr_vdw(const Atoms& at, FFDev ff, const int* row_ptr, const int* nbr_j, const int* nbr_s, EnergyAcc acc,
                  hipStream_t s);
// Coulomb and QEq self energy (after QEq): forces into F_nb.
hipError_t xr_nonbonded(const Atoms& at, FFDev ff, const int* row_ptr, const int* nbr_j, const int* nbr_s,
                        EnergyAcc acc, double2* dEdq, hipStream_t s);
// Adds the charge-response term (full derivative through QEq) to forces.
hipError_t xr_charge_response(const Atoms& at, FFDev ff, const int* row_ptr, const int* nbr_j, const int* nbr_s,
                              const double* lam, double* forces, hipStream_t s);
// Per-system channel sums [ns][NCH] from the per-atom channels.
hipError_t xr_energy_reduce(const Systems& S, EnergyAcc acc, double* chan, hipStream_t s);

// The reported components from one system's channel sums.
inline void xr_components(const double* ch, double* out) {
  for (int k = 0; k < NCOMP; ++k) out[k] = 0.0;
  out[C_BOND] = ch[CH_BOND] + ch[CH_TRIPLE];
  out[C_ATOM] = ch[CH_ATOM];
  out[C_LONE] = ch[CH_LP] + ch[CH_C2];
  out[C_ANGLE] = ch[CH_ANGLE];
  out[C_PENALTY] = ch[CH_PEN];
  out[C_COA] = ch[CH_COA];
  out[C_HBOND] = ch[CH_HB];
  out[C_TORSION] = ch[CH_TOR];
  out[C_CONJ] = ch[CH_CONJ];
  out[C_VDW] = ch[CH_VDW];
  out[C_COULOMB] = ch[CH_COUL];
  out[C_QEQ] = ch[CH_QEQ];
}
