// QEq launch interface (qeq.hip).
#pragma once
#include "common.h"

struct CgArgs {
  int n;
  const int* type;
  FFDev ff;
  const int* row_ptr;
  const int* nbr_j;
  const double* hval;
  double2 *X, *R, *AP;
  double4 *ZP0, *ZP1;  // (z_s, z_t, p_s, p_t), alternating by iteration
  double* inv_diag;
  double* partials;  // xr_cg_partials(n) doubles
  void* state;       // xr_cg_state_bytes() bytes
  double tol;        // relative residual per right-hand side
  int max_iter;
  int first_check;   // iterations before the first convergence read-back
  int check_every;   // iterations between later read-backs
  double total_charge;
  double* q;         // out: charges
  double* mu;        // out: 1 double
  const unsigned* abort_flag;  // device word; nonzero stops the solve (input errors)
};

hipError_t xr_edge_prep(int n, const double* x, const int* type, CellDev cell, FFDev ff, const int* row_ptr,
                        const int* nbr_j, const int* nbr_s, double bond_cut, double* hval, unsigned char* bflag,
                        int* bond_count, hipStream_t s);
hipError_t xr_bond_fill(int n, const int* row_ptr, const unsigned char* bflag, const int* bond_ptr, int cap,
                        int* bond_edge, int* bond_i, hipStream_t s);
size_t xr_cg_partials(int n);
size_t xr_cg_state_bytes();
hipError_t xr_cg_solve(const CgArgs& a, hipStream_t s, int* iterations);
