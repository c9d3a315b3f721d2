// QEq launch interface (qeq.hip).
#pragma once
#include "common.h"

// CG_REFRESH: the system's next direction launch replaces the recurrence
// residual by the FP64 true residual b - A x (and z, p = 0), the update
// launch after it reduces the new residual, and the recurrence restarts.
enum CgState { CG_RUNNING = 0, CG_DONE = 1, CG_REFRESH = 2, CG_FAILED = 3, CG_STALLED = 4 };

// Per-system solver state (device).
struct CgSys {
  int done;       // CgState
  int iters;      // iterations executed by this system
  int k0;         // global iteration index of the last (re)start (set by the refresh)
  int restarts;
  int kr;         // CG_REFRESH: launch whose direction kernel computes the true residual
  double rr0[2];  // true squared residuals at the last restart (mixed precision)
};

struct CgCounters {
  int settled;  // systems in a final state (done, failed, stalled)
  int abort;    // copy of the upstream overflow word, taken at init
};

// The state block holds the counters, then CgSys[ns] at byte offset 64.
inline CgSys* xr_cg_systems(void* state) { return reinterpret_cast<CgSys*>(static_cast<char*>(state) + 64); }

struct CgArgs {
  Systems S;
  int nd, nu;          // total direction / update blocks over all systems
  const int* type;
  FFDev ff;
  const int* row_ptr;
  const int* nbr_j;
  const double* hval;
  const float* hval32;  // non-null: mixed precision (FP32 A p sweeps, FP64 restarts)
  double2 *X, *R, *AP;
  double4 *ZP0, *ZP1;  // (z_s, z_t, p_s, p_t), alternating by iteration
  float4 *ZQ0, *ZQ1;   // FP32 copies gathered by the mixed-precision sweep
  const unsigned char* join;  // per atom: strong bonds to the previous three atoms (edge pass)
  int n;                      // atoms over all systems
  double* minv;               // block-Jacobi inverse rows, column-major: entry c of row i at minv[c * n + i]
  int2* binfo;                // per atom: (local row of its block start, block size)
  double* eta;
  double* partials;  // xr_cg_partials(nd, nu) doubles
  void* state;       // xr_cg_state_bytes(ns) bytes
  double tol;        // relative residual per right-hand side
  int max_iter;      // per system
  int first_check;   // iterations before the first read-back
  int check_every;   // iterations between later read-backs
  const double2* rhs;    // optional right-hand sides per atom (default (-chi, -1))
  bool neutral;          // normalise the result to zero sum instead of the systems' total charges
  const double2* guess;  // optional starting (s, t) per atom
  double2* keep;         // optional: receives the solution (s, t) per atom
  double* q;             // out: charges
  double* mu;            // out: one double per system
  const unsigned* abort_flag;  // device word; nonzero stops the solve (buffer overflow)
  int* poll_host;              // pinned host memory, two slots of two ints (settled, abort)
  hipEvent_t poll_ev[2];       // completion of each slot's copy
};

struct CgOutcome {
  bool aborted;    // stopped on abort_flag
  bool exhausted;  // launch bound reached before every system settled
  int sweeps;      // iterations launched
};

hipError_t xr_edge_prep(int n, const double* x, const int* type, const Systems& S, FFDev ff, const int* row_ptr,
                        const int* nbr_j, const int* nbr_s, double bond_cut, double* hval, float* hval32,
                        unsigned char* bflag, int* bond_count, unsigned char* join, hipStream_t s);
hipError_t xr_bond_fill(int n, const int* row_ptr, const unsigned char* bflag, const int* bond_ptr, int cap,
                        int* bond_edge, int* bond_i, hipStream_t s);
void xr_cg_blocks(int n, int* ndir, int* nupd);
size_t xr_cg_partials(int nd, int nu);
size_t xr_cg_state_bytes(int ns);
size_t xr_cg_block_doubles(int n);  // doubles of the block-Jacobi inverse rows of n atoms
hipError_t xr_cg_solve(const CgArgs& a, hipStream_t s, CgOutcome* out);
