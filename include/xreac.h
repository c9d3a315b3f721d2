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
 * thread-safe with respect to each other.
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
 * pass, CG direction+A p sweep, CG update, nonbonded, bond-order forward,
 * many-body forward, adjoint. */
enum {
  XREAC_TIMER_NEIGHBOR_COUNT = 0,
  XREAC_TIMER_NEIGHBOR_FILL = 1,
  XREAC_TIMER_EDGE = 2,
  XREAC_TIMER_CG_DIRECTION = 3,
  XREAC_TIMER_CG_UPDATE = 4,
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
} xreac_system;

typedef struct {
  double qeq_tolerance;        /* relative residual of each CG right-hand side */
  int32_t qeq_max_iterations;
  int32_t lammps_angles;       /* 1: LAMMPS near-collinear angle derivative (fixed-charge mode) */
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

/* Kernel timing with hipEvents on the context stream (no reference
 * counterpart; measurement only).  timer_mask bit k enables timer k (0: off);
 * every launch of an enabled kernel records start/stop events with its
 * dispatch and its elapsed time is summed per timer. */
int xreac_profile_enable(xreac_ctx* ctx, int timer_mask);
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
