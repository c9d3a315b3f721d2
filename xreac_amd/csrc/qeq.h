// QEq launch interface (qeq.hip).
#pragma once
#include "common.h"

// CG_REFRESH: the system's next launch replaces the recurrence residual by
// the FP64 true residual r = b - A x (and u = M^-1 r), the launch after it
// computes w = A u, and the recurrence restarts from there.
enum CgState { CG_RUNNING = 0, CG_DONE = 1, CG_REFRESH = 2, CG_FAILED = 3, CG_STALLED = 4 };

// Solver control of one system as one launch leaves it for the next.
struct CgCtl {
  int done;       // CgState
  int k0;         // launch of the last (re)start's first iteration
  int kr;         // CG_REFRESH: launch that computes the true residual
  int restarts;
  int it;         // iterations executed
  double rr0[2];  // squared residuals at the last (re)start (mixed precision)
  double gam[2];  // r.u of the last iteration, per right-hand side
  double alp[2];  // step of the last iteration
};

// Per-system solver state (device).  Launch k reads ctl[(k - 1) & 1] and
// its first block writes ctl[k & 1], so no block reads a slot that another
// block of the same launch writes (final states, which stop every block, go
// to both).
struct CgSys {
  CgCtl ctl[2];
  int iters;  // iterations when the system settled
  int ks;     // launch in which it settled
};

struct CgCounters {
  int settled;  // systems in a final state (done, failed, stalled)
  int abort;    // copy of the upstream overflow word, taken at init
};

// The state block holds the counters, then CgSys[ns] at byte offset 64.
inline CgSys* xr_cg_systems(void* state) { return reinterpret_cast<CgSys*>(static_cast<char*>(state) + 64); }

// CG vectors, (s, t) pairs per atom (pipelined CG recurrences).
struct CgVec {
  double2 *X, *R, *U, *W, *Z, *Q, *S, *P;
  double2* M[2];   // M^-1 w, alternating by launch (gathered by the next launch)
  float2* M32[2];  // FP32 copies gathered by the mixed-precision sweep
};

// Per CG iteration block: its descriptor (system, atom range [y, z) of its
// first chunk, that chunk's system-local index) and the row pointers of that
// chunk's rows, rp[t] = row_ptr[y + min(t, z - y)] for t <= CG_CHUNK_ROWS_MAX,
// so a wave of the block finds its rows' entry ranges with one load indexed by
// the block, without waiting for the descriptor.
constexpr int CG_CHUNK_ROWS_MAX = 64;  // rows of one chunk at most
struct CgBlock {
  int4 desc;
  int rp[CG_CHUNK_ROWS_MAX + 12];  // 65 used; 320 bytes per block
};

struct CgArgs {
  Systems S;
  int nd, nu;          // total iteration blocks / chunks over all systems
  int nd_small;        // iteration blocks of small systems, numbered first
  const int* type;
  FFDev ff;
  const int* row_ptr;
  const int* nbr_j;
  const unsigned short* col16;  // nbr_j minus the system's first atom, 2 bytes (systems of at most 65,536 atoms) or null
  const double* hval;
  const float* hval32;  // non-null: mixed precision (FP32 A m sweeps, FP64 restarts)
  CgVec v;
  const unsigned char* join;  // per atom: strong bonds to the previous three atoms (edge pass)
  int n;                      // atoms over all systems
  double* minv;               // block-Jacobi inverse rows, column-major: entry c of row i at minv[c * n + i]
  int2* binfo;                // per atom: (chunk row of its block start, block size)
  int2* chunks;               // per chunk: atom range [x, y)
  CgBlock* blocks;            // per iteration block: descriptor and first chunk's row pointers
  double* eta;
  double* partials;  // xr_cg_partials(nd, nu) doubles
  void* state;       // xr_cg_state_bytes(ns) bytes
  double tol;        // relative residual per right-hand side
  int max_iter;      // per system
  int first_check;   // launches before the first read-back
  int check_every;   // launches between later read-backs
  int fixed_launches;  // > 0: exactly this many iteration launches and no read-back (a recorded solve)
  const double2* rhs;    // optional right-hand sides per atom (default (-chi, -1))
  bool neutral;          // normalise the result to zero sum instead of the systems' total charges
  const double2* guess;  // optional starting (s, t) per atom
  double2* keep;         // optional: receives the solution (s, t) per atom
  double* q;             // out: charges
  double* mu;            // out: one double per system
  const unsigned* abort_flag;  // device word; nonzero stops the solve (buffer overflow)
  int* poll_host;              // pinned host memory, two slots of two ints (settled, abort)
  hipEvent_t poll_ev[2];       // completion of each slot's copy
  // Optional host call made once the solve's first few launches are queued and
  // before the first wait: the caller issues independent work to other streams
  // there, so the solve starts without waiting for those launches.
  hipError_t (*after_start)(void*);
  void* after_start_arg;
  // Optional work that follows the solve on the same stream.  When given, it is
  // issued (after the charges) behind the first chunk of launches, before the
  // first wait; CgOutcome::tail_done says whether that issue stands or more
  // launches followed it, in which case the caller issues the work again.
  hipError_t (*tail)(void*);
  void* tail_arg;
};

struct CgOutcome {
  bool aborted;    // stopped on abort_flag
  bool exhausted;  // launch bound reached before every system settled
  int sweeps;      // iteration launches issued
  bool tail_done;  // the tail issued behind the first chunk is final
};

// slab: the fill's sorted rows (row i at slab[i * slab_stride]), moved to nbr_j / nbr_s on the way; null
// when nbr_j / nbr_s already hold the rows (a supplied list).  col16 (or null): receives the
// neighbor indices relative to the system's first atom as 2-byte columns for the CG sweeps (every
// system of at most 65,536 atoms).  hb_cand (one int2 slot per edge) receives each row's H-bond
// candidates as (j, shift code) at the start of the row's range, hb_count their number per row.
hipError_t xr_edge_prep(int n, const double* x, const int* type, const Systems& S, FFDev ff, const int* row_ptr,
                        const void* slab, int slab_stride, int* nbr_j, int* nbr_s, unsigned short* col16,
                        double bond_cut, double* hval, float* hval32, unsigned char* bflag, int* bond_count,
                        unsigned char* join, int2* hb_cand, int* hb_count, hipStream_t s);
hipError_t xr_bond_fill(int n, const int* row_ptr, const unsigned char* bflag, const int* bond_ptr, int cap,
                        int* bond_edge, int* bond_i, hipStream_t s);
// Iteration blocks (ndir), chunks (nupd) and nominal rows per chunk of a
// system of n atoms; small systems (xr_cg_small) take the 512-thread shape.
void xr_cg_blocks(int n, int* ndir, int* nupd, int* chunk);
bool xr_cg_small(int n);
size_t xr_cg_partials(int nd, int nu);
size_t xr_cg_state_bytes(int ns);
size_t xr_cg_block_doubles(int n);  // doubles of the block-Jacobi inverse rows of n atoms
hipError_t xr_cg_solve(const CgArgs& a, hipStream_t s, CgOutcome* out);
