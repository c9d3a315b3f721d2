/*
 * xreac C ABI - ReaxFF energy, QEq charges, fixed-charge forces and bond
 * properties on one AMD Instinct MI355X (gfx950).
 *
 * This is the drop-in boundary for the evaluation path of hsulab/xreac 0.10.0:
 *   ForceField tables  -> xreac_ff_create         (replaces the parsed
 *                         ForceField consumed by EnergyModel, energy.py:65-87)
 *   Calculator.evaluate -> xreac_evaluate          (calculator.py:76-165)
 *   Neighbors (i, j, S) -> xreac_copy_neighbors    (neighbors.py:101-131)
 *   bond-order state    -> xreac_copy_bonds        (energy.py:163-178)
 * Plain pointers and sizes only.  Arrays marked "device" live in GPU memory
 * (hipMalloc or a framework allocator); arrays marked "host" in CPU memory.
 * All calls on one context are issued on the context's stream and are not
 * thread-safe with respect to each other.  Calls on different contexts may
 * come from different threads; the library runs one evaluation (including a
 * recorded step's replay) at a time per device and returns to the other
 * threads' calls as each finishes, so such calls overlap only in host work
 * and in the integrator kernels.
 */
#ifndef XREAC_H
#define XREAC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  XREAC_INVALID carries a message equal to the ValueError text
 * the reference raises for the same condition. */
enum {
  XREAC_OK = 0,
  XREAC_INVALID = 1,
  XREAC_DEVICE = 2,
  XREAC_NOMEM = 3
};

/* Column layouts of the per-type and per-pair tables. */
enum {
  XREAC_ATOM_PARAMS = 20, /* r_s r_pi r_pi_pi valency valency_e valency_boc valency_val mass chi
                             eta gamma p_lp2 p_ovun2 p_ovun5 p_val3 p_val5 p_hbond is_carbon
                             mass==12.0 mass==15.999  (eta already doubled) */
  XREAC_PAIR_PARAMS = 28, /* p_bo1..p_bo6 r_s r_p r_pp ovc v13cor p_boc3 p_boc4 p_boc5 De_s De_p
                             De_pp p_be1 p_be2 p_ovun1 D alpha gamma_w r_vdW rcore ecore acore
                             gamma */
  XREAC_COMPONENTS = 14   /* bond atom lone_pair molecule angle penalty angle_conjugation
                             hydrogen_bond torsion conjugation vdw coulomb electric_field qeq */
};

/* Kernel timers (xreac_profile_read): neighbor count, neighbor fill, edge
 * pass, CG iteration (A m sweep and vector recurrences), CG set-up (chunks and
 * block-Jacobi inverse), nonbonded, bond-order forward, many-body forward,
 * adjoint. */
enum {
  XREAC_TIMER_NEIGHBOR_COUNT = 0,
  XREAC_TIMER_NEIGHBOR_FILL = 1,
  XREAC_TIMER_EDGE = 2,
  XREAC_TIMER_CG_ITERATION = 3,
  XREAC_TIMER_CG_INIT = 4,
  XREAC_TIMER_NONBONDED = 5,
  XREAC_TIMER_BOND_FORWARD = 6,
  XREAC_TIMER_MANYBODY = 7,
  XREAC_TIMER_BACKWARD = 8,
  XREAC_TIMERS = 9,
  XREAC_TIMER_ALL = (1 << 9) - 1
};

typedef struct xreac_ff xreac_ff;
typedef struct xreac_ctx xreac_ctx;

/* Parameter tables (host memory, copied to the device by xreac_ff_create). */
typedef struct {
  int32_t ntypes;
  int32_t vdw_type;              /* 1 shielded, 2 inner wall, 3 both */
  const double* general;         /* [39] */
  const double* atom;            /* [ntypes][XREAC_ATOM_PARAMS] */
  const double* pair;            /* [ntypes*ntypes][XREAC_PAIR_PARAMS] */
  const int32_t* angle_offsets;  /* [ntypes^3 + 1], key (t_i*nt + t_centre)*nt + t_k */
  const double* angle_rows;      /* [angle_offsets[ntypes^3]][7] rows with |v1| > 0.001 */
  const double* torsion;         /* [ntypes^4][5], key ((t_i*nt + t_j)*nt + t_k)*nt + t_l */
  const uint8_t* torsion_on;     /* [ntypes^4] 1 if present and active */
  const double* hbond;           /* [ntypes^3][4], key (t_donor*nt + t_H)*nt + t_acceptor */
  const uint8_t* hbond_on;       /* [ntypes^3] 1 if present with r0 > 0 */
} xreac_ff_tables;

/* One system. */
typedef struct {
  int32_t n;
  const double* positions;   /* device [n*3], Angstrom, any image branch */
  const int32_t* types;      /* device [n], indices into the force-field tables */
  const double* cell;        /* host [9], lattice vectors as rows; NULL if nonperiodic */
  int32_t pbc[3];
  double total_charge;       /* e */
  /* Optional caller-supplied full directed neighbor list (calculator.py:113-115,
   * neighbors.py:101-131), device memory in canonical (i, j, Sx, Sy, Sz) order as
   * CSR: neighbor_row_ptr [n+1], neighbor_j [M], neighbor_shifts [M*3]; NULL row
   * pointer: the list is built on the device.  Edges beyond the cutoff (a skin)
   * are allowed and contribute nothing. */
  const int32_t* neighbor_row_ptr;
  const int32_t* neighbor_j;
  const int32_t* neighbor_shifts;
  int64_t neighbor_count;
} xreac_system;

enum {
  XREAC_PRECISION_FP64 = 0,    /* FP64 CG throughout */
  XREAC_PRECISION_MIXED = 1    /* FP32 A m sweeps, FP64 vectors and dots, FP64 true-residual restarts */
};

typedef struct {
  double qeq_tolerance;        /* relative residual of each CG right-hand side */
  int32_t qeq_max_iterations;  /* per system */
  int32_t lammps_angles;       /* 1: LAMMPS near-collinear angle derivative (fixed-charge mode) */
  int32_t precision;           /* XREAC_PRECISION_* */
  int32_t warm_start;          /* 1: start the CG from result->qeq_state (batched entry point) */
  int32_t full_derivative;     /* 1: charge-response forces (derivative through QEq, one adjoint solve);
                                  the plain angle derivative is then used, as calculator.py:79-86 does */
  int32_t fixed_charges;       /* 1: no QEq; the charges array is read as the atomic charges (the
                                  EnergyModel.components(x, fixed_charges) hook, energy.py:180-219) */
  double neighbor_skin;        /* > 0: ASE's per-atom displacement allowance (Angstrom), as the
                                  reference's ASE calculator uses it (ase.py:39-40, 90-93, 107-131):
                                  keep the device-built list within rc + 2 skin and reuse it on later
                                  calls of the same shape (same force field, atom offsets, cells and
                                  pbc) until an atom has moved more than skin or a type changed.  Each
                                  call filters the kept rows to the cutoff, so results are bitwise those
                                  of a fresh build.  Ignored with a supplied list; 0 (the default): a
                                  fresh list every call */
} xreac_options;

typedef struct {
  double components[XREAC_COMPONENTS]; /* host, kcal/mol */
  double* forces;              /* device [n*3], kcal/mol/Angstrom */
  double* charges;             /* device [n], e */
  double* total_bond_orders;   /* device [n] or NULL */
  double* lone_pairs;          /* device [n] or NULL */
  int32_t* bond_counts;        /* device [n] or NULL */
  int64_t num_edges;           /* out: directed neighbor edges */
  int64_t num_bonds;           /* out: directed active bonds */
  int32_t qeq_iterations;      /* out */
} xreac_result;

/* A batch of independent systems packed atom-wise.  The reference evaluates
 * one system per call; this is the same evaluation for many at once.
 * System s owns atoms [atom_offsets[s], atom_offsets[s+1]) of every per-atom
 * array; its results are bitwise those of evaluating it alone. */
typedef struct {
  int32_t num_systems;
  const int32_t* atom_offsets;  /* host [num_systems + 1], atom_offsets[0] = 0, nondecreasing */
  const double* positions;      /* device [N*3], N = atom_offsets[num_systems] */
  const int32_t* types;         /* device [N] */
  const double* cells;          /* host [num_systems*9] (rows = lattice vectors), or NULL: all nonperiodic */
  const int32_t* pbc;           /* host [num_systems*3], or NULL: periodic along all three cell vectors */
  const double* total_charges;  /* host [num_systems], or NULL: all neutral */
  /* Optional supplied neighbor lists of all systems as one CSR over the packed
   * atoms (global indices), as in xreac_system; NULL: built on the device. */
  const int32_t* neighbor_row_ptr;
  const int32_t* neighbor_j;
  const int32_t* neighbor_shifts;
  int64_t neighbor_count;
} xreac_batch;

typedef struct {
  double* components;          /* host [num_systems][XREAC_COMPONENTS], kcal/mol */
  int32_t* status;             /* host [num_systems]: XREAC_OK, or XREAC_INVALID (xreac_batch_error) */
  int32_t* qeq_iterations;     /* host [num_systems] */
  double* forces;              /* device [N*3] */
  double* charges;             /* device [N] */
  double* total_bond_orders;   /* device [N] or NULL */
  double* lone_pairs;          /* device [N] or NULL */
  int32_t* bond_counts;        /* device [N] or NULL */
  double* qeq_state;           /* device [N*2] or NULL: (s, t) per atom; the starting guess when
                                  options.warm_start is set; overwritten with the solution */
  int64_t num_edges;           /* out: directed neighbor edges, all systems */
  int64_t num_bonds;           /* out: directed active bonds, all systems */
} xreac_batch_result;

/* Force field. */
int xreac_ff_create(const xreac_ff_tables* tables, xreac_ff** out);
void xreac_ff_destroy(xreac_ff* ff);

/* Context: one per device/stream; owns grow-only device workspace.
 * stream is a hipStream_t (NULL = default stream). */
int xreac_ctx_create(int device, void* stream, xreac_ctx** out);
void xreac_ctx_destroy(xreac_ctx* ctx);
const char* xreac_error_message(const xreac_ctx* ctx);

/* Energy, QEq charges and fixed-charge forces (calculator.py:76-165). */
int xreac_evaluate(xreac_ctx* ctx, const xreac_ff* ff, const xreac_system* sys, const xreac_options* opt,
                   xreac_result* out);

/* Bytes of the static device scratch array behind xreac_evaluate_static
 * (64 MiB per device, held from the library's first use on a device until
 * the process ends). */
#define XREAC_STATIC_SCRATCH_BYTES (64u << 20)

/* xreac_evaluate with no allocation of any kind: no counterpart in the
 * reference.  It is for callers that cannot allocate device or pinned memory
 * mid-run (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Every workspace comes from a file-scope static device
 * array of XREAC_STATIC_SCRATCH_BYTES; the side stream, events and pinned
 * read-back slots are made by xreac_ctx_create.  A system whose workspace
 * (about 18 KB per atom of liquid water, i.e. up to roughly 3,600 such atoms)
 * does not fit returns XREAC_NOMEM with a message naming the size.  Results
 * are bitwise those of xreac_evaluate.  All calls on one device share the
 * array: calls that overlap in time on one device (two contexts, two host
 * threads) race; the last call's lists (xreac_copy_neighbors,
 * xreac_copy_bonds) stay readable until the next call on that device. */
int xreac_evaluate_static(xreac_ctx* ctx, const xreac_ff* ff, const xreac_system* sys, const xreac_options* opt,
                          xreac_result* out);

/* The same evaluation over a batch (calculator.py:76-165 per system).  Returns
 * XREAC_OK when the batch ran; each system then has its own status, and a
 * failed system (e.g. coincident atoms) does not affect the others.  Other
 * return codes mean nothing was evaluated. */
int xreac_evaluate_batch(xreac_ctx* ctx, const xreac_ff* ff, const xreac_batch* batch, const xreac_options* opt,
                         xreac_batch_result* out);

/* Message of system s of the last batch ("" if it succeeded). */
const char* xreac_batch_error(const xreac_ctx* ctx, int32_t system);

/* Skin lists built on this context so far, one per system and rebuild
 * (the reference's ReaxFFCalculator.neighbor_list_builds, ase.py:130). */
int64_t xreac_neighbor_list_builds(const xreac_ctx* ctx);

/* Velocity Verlet over a packed batch, one step in two stages around a force
 * evaluation (no reference counterpart: the reference has no dynamics).
 * Units A, fs, amu (force-field masses), kcal/mol.  All pointers are device
 * memory: offsets [num_systems + 1], types [N], dt_fs [num_systems] (0 keeps
 * a system still), positions/velocities/forces [N*3].
 *   stage 0: v += dt/2 F/m, then x += dt v   (forces of the current positions)
 *   stage 1: v += dt/2 F/m                   (forces of the new positions);
 *            kinetic [num_systems] (or NULL) receives sum m v^2 / 2 in kcal/mol. */
int xreac_verlet(xreac_ctx* ctx, const xreac_ff* ff, int32_t num_systems, const int32_t* offsets,
                 const int32_t* types, const double* dt_fs, double* positions, double* velocities,
                 const double* forces, int32_t stage, double* kinetic);

/* Starting guesses for the QEq solves of the next velocity-Verlet step (no
 * reference counterpart: the reference has no dynamics).  Per system s of the
 * packed batch, with k = order[s] (clamped to 0..3), the (s, t) state of each
 * atom is replaced by the degree-k polynomial through its current value and
 * the k newest of the three previous solutions in history, evaluated one
 * equal time step ahead (k = 0 keeps it); history then shifts by one, the
 * current value becoming its newest entry.  Use it between solves of systems
 * stepped with a fixed dt, k at most the number of earlier solutions.  All
 * pointers are device memory: offsets [num_systems + 1], order
 * [num_systems], qeq_state [num_atoms*2], history [3][num_atoms*2]. */
int xreac_qeq_extrapolate(xreac_ctx* ctx, int32_t num_systems, int32_t num_atoms, const int32_t* offsets,
                          const int32_t* order, double* qeq_state, double* history);

/* One FIRE iteration per system of a packed batch, as the reference's native
 * relaxation loop (calculator.py:213-246): if max |F| <= force_tolerance the
 * system is marked converged and not moved; otherwise velocities are mixed
 * and kicked and positions moved (largest atom step capped at 0.1 A).
 * Device memory: offsets [num_systems + 1], positions/velocities/forces [N*3],
 * state [num_systems*4] = (dt, alpha, positive steps, iteration), start
 * (0.02, 0.1, 0, 0) with zero velocities; status [num_systems]: 0 running,
 * 1 converged, 2 iteration limit reached (systems with status != 0 are left
 * untouched). */
int xreac_fire_step(xreac_ctx* ctx, int32_t num_systems, const int32_t* offsets, double* positions,
                    double* velocities, const double* forces, double* state, int32_t* status,
                    double force_tolerance, int32_t max_iterations);

/* Integrator stages of a recorded step (xreac_step_create), run in this order
 * around its evaluation: velocity-Verlet stage 0, QEq extrapolation, the
 * evaluation, velocity-Verlet stage 1, one FIRE iteration.  A NULL dt_fs,
 * qeq_order or fire_state leaves that stage out.  Arrays as in xreac_verlet,
 * xreac_qeq_extrapolate and xreac_fire_step (device memory); stage 0 reads the
 * forces array's content (the previous step's forces). */
typedef struct {
  const double* dt_fs;        /* [num_systems]; 0 keeps a system still */
  double* velocities;         /* [N*3] */
  double* kinetic;            /* [num_systems] or NULL */
  const int32_t* qeq_order;   /* [num_systems] */
  double* qeq_history;        /* [3][N*2] */
  double* fire_state;         /* [num_systems*4] */
  int32_t* fire_status;       /* [num_systems] */
  double fire_force_tolerance;
  int32_t fire_max_iterations;
} xreac_step_stages;

typedef struct xreac_step xreac_step;

/* A step of a fixed batch recorded once as a HIP graph and then replayed: no
 * reference counterpart (the reference evaluates one system per call, host
 * driven).  It is for serving and dynamics loops that step the same systems
 * many times: a replay is one graph launch and one wait instead of about a
 * hundred host-issued launches and the QEq convergence read-backs.  The
 * batch's host arrays are copied; its device arrays, the result's device
 * arrays and the stages' arrays are used in place by every run, so their
 * contents may change between runs (positions, dt, extrapolation orders) but
 * not their addresses or sizes.  The step owns its workspaces.  The context
 * needs its own stream (xreac_ctx_create with a non-NULL stream) and must
 * outlive the step.  Neither supplied neighbor lists nor neighbor_skin. */
int xreac_step_create(xreac_ctx* ctx, const xreac_ff* ff, const xreac_batch* batch, const xreac_options* opt,
                      const xreac_step_stages* stages, const xreac_batch_result* result, xreac_step** out);
/* Run the step once and wait for it; the host arrays of result (components,
 * status, qeq_iterations) receive the evaluation's, as xreac_evaluate_batch.
 * The first run executes the step unrecorded and records it.  A replay whose
 * edge or bond buffers overflowed, or whose QEq needed more launches than were
 * recorded, puts back every array the step modifies and runs the step
 * unrecorded, then records it again; either way the results are bitwise those
 * of the unrecorded step. */
int xreac_step_run(xreac_step* step, xreac_batch_result* result);
/* Replays, fallbacks to an unrecorded run after a replay, and recordings so far. */
int xreac_step_stats(const xreac_step* step, int64_t* replays, int64_t* fallbacks, int64_t* recordings);
/* Frees the step's graph and workspaces; it does not use the context. */
void xreac_step_destroy(xreac_step* step);

/* Kernel timing with hipEvents on the launching stream (no reference
 * counterpart; measurement only).  timer_mask bit k enables timer k (0: off);
 * each sampled launch of an enabled kernel records start/stop events with
 * its dispatch and its elapsed time is summed per timer. */
int xreac_profile_enable(xreac_ctx* ctx, int timer_mask);
/* Record events on one dispatch in `every` of each enabled timer (default 1),
 * so a timed run keeps the event overhead small.  No reference counterpart
 * (measurement only). */
int xreac_profile_stride(xreac_ctx* ctx, int every);
int xreac_profile_read(xreac_ctx* ctx, int timer, double* total_ms, int64_t* launches, int reset);

/* Canonical neighbor list of the last evaluation (host arrays): row_ptr [n+1],
 * j [M], shifts [M*3].  capacity is the length of j; returns XREAC_INVALID if
 * it is too small. */
int xreac_copy_neighbors(xreac_ctx* ctx, int32_t* row_ptr, int32_t* j, int32_t* shifts, int64_t capacity);

/* Active bonds of the last evaluation (host arrays of length num_bonds):
 * row atom, far atom, image shift [*3], corrected bond order BO. */
int xreac_copy_bonds(xreac_ctx* ctx, int32_t* i, int32_t* j, int32_t* shifts, double* bo, int64_t capacity);

#ifdef __cplusplus
}
#endif
#endif /* XREAC_H */
