"""ctypes binding of the HIP library ``libxreac_hip.so`` (C ABI in include/xreac.h).

The library is built in-tree by ``xreac_amd.build``.  There is no fallback: if
the library is missing or fails to load, every entry point raises.
"""

from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libxreac_hip.so")

XREAC_OK, XREAC_INVALID, XREAC_DEVICE, XREAC_NOMEM = 0, 1, 2, 3
PRECISION_FP64, PRECISION_MIXED = 0, 1
NCOMP = 14
STATIC_SCRATCH_BYTES = 64 << 20  # XREAC_STATIC_SCRATCH_BYTES

EXPORTED = (
    "xreac_ff_create",
    "xreac_ff_destroy",
    "xreac_ctx_create",
    "xreac_ctx_destroy",
    "xreac_error_message",
    "xreac_evaluate",
    "xreac_evaluate_static",
    "xreac_evaluate_batch",
    "xreac_batch_error",
    "xreac_verlet",
    "xreac_fire_step",
    "xreac_qeq_extrapolate",
    "xreac_neighbor_list_builds",
    "xreac_copy_neighbors",
    "xreac_copy_bonds",
    "xreac_profile_enable",
    "xreac_profile_stride",
    "xreac_profile_read",
    "xreac_step_create",
    "xreac_step_run",
    "xreac_step_stats",
    "xreac_step_destroy",
)
TIMERS = (
    "neighbor_count",
    "neighbor_fill",
    "edge",
    "cg_iteration",
    "cg_init",
    "nonbonded",
    "bond_forward",
    "manybody",
    "backward",
)


class FFTables(C.Structure):
    _fields_ = [
        ("ntypes", C.c_int32),
        ("vdw_type", C.c_int32),
        ("general", C.c_void_p),
        ("atom", C.c_void_p),
        ("pair", C.c_void_p),
        ("angle_offsets", C.c_void_p),
        ("angle_rows", C.c_void_p),
        ("torsion", C.c_void_p),
        ("torsion_on", C.c_void_p),
        ("hbond", C.c_void_p),
        ("hbond_on", C.c_void_p),
    ]


class System(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("positions", C.c_void_p),
        ("types", C.c_void_p),
        ("cell", C.c_void_p),
        ("pbc", C.c_int32 * 3),
        ("total_charge", C.c_double),
        ("neighbor_row_ptr", C.c_void_p),
        ("neighbor_j", C.c_void_p),
        ("neighbor_shifts", C.c_void_p),
        ("neighbor_count", C.c_int64),
    ]


class Options(C.Structure):
    _fields_ = [
        ("qeq_tolerance", C.c_double),
        ("qeq_max_iterations", C.c_int32),
        ("lammps_angles", C.c_int32),
        ("precision", C.c_int32),
        ("warm_start", C.c_int32),
        ("full_derivative", C.c_int32),
        ("fixed_charges", C.c_int32),
        ("neighbor_skin", C.c_double),
    ]


class Batch(C.Structure):
    _fields_ = [
        ("num_systems", C.c_int32),
        ("atom_offsets", C.c_void_p),
        ("positions", C.c_void_p),
        ("types", C.c_void_p),
        ("cells", C.c_void_p),
        ("pbc", C.c_void_p),
        ("total_charges", C.c_void_p),
        ("neighbor_row_ptr", C.c_void_p),
        ("neighbor_j", C.c_void_p),
        ("neighbor_shifts", C.c_void_p),
        ("neighbor_count", C.c_int64),
    ]


class BatchResult(C.Structure):
    _fields_ = [
        ("components", C.c_void_p),
        ("status", C.c_void_p),
        ("qeq_iterations", C.c_void_p),
        ("forces", C.c_void_p),
        ("charges", C.c_void_p),
        ("total_bond_orders", C.c_void_p),
        ("lone_pairs", C.c_void_p),
        ("bond_counts", C.c_void_p),
        ("qeq_state", C.c_void_p),
        ("num_edges", C.c_int64),
        ("num_bonds", C.c_int64),
    ]


class StepStages(C.Structure):
    _fields_ = [
        ("dt_fs", C.c_void_p),
        ("velocities", C.c_void_p),
        ("kinetic", C.c_void_p),
        ("qeq_order", C.c_void_p),
        ("qeq_history", C.c_void_p),
        ("fire_state", C.c_void_p),
        ("fire_status", C.c_void_p),
        ("fire_force_tolerance", C.c_double),
        ("fire_max_iterations", C.c_int32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("components", C.c_double * NCOMP),
        ("forces", C.c_void_p),
        ("charges", C.c_void_p),
        ("total_bond_orders", C.c_void_p),
        ("lone_pairs", C.c_void_p),
        ("bond_counts", C.c_void_p),
        ("num_edges", C.c_int64),
        ("num_bonds", C.c_int64),
        ("qeq_iterations", C.c_int32),
    ]


class LibraryMissing(RuntimeError):
    pass


_LIB = None


def load():
    """Load and type the library once; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    lib.xreac_ff_create.argtypes = [C.POINTER(FFTables), C.POINTER(vp)]
    lib.xreac_ff_create.restype = C.c_int
    lib.xreac_ff_destroy.argtypes = [vp]
    lib.xreac_ff_destroy.restype = None
    lib.xreac_ctx_create.argtypes = [C.c_int, vp, C.POINTER(vp)]
    lib.xreac_ctx_create.restype = C.c_int
    lib.xreac_ctx_destroy.argtypes = [vp]
    lib.xreac_ctx_destroy.restype = None
    lib.xreac_error_message.argtypes = [vp]
    lib.xreac_error_message.restype = C.c_char_p
    lib.xreac_evaluate.argtypes = [vp, vp, C.POINTER(System), C.POINTER(Options), C.POINTER(Result)]
    lib.xreac_evaluate.restype = C.c_int
    lib.xreac_evaluate_static.argtypes = [vp, vp, C.POINTER(System), C.POINTER(Options), C.POINTER(Result)]
    lib.xreac_evaluate_static.restype = C.c_int
    lib.xreac_evaluate_batch.argtypes = [vp, vp, C.POINTER(Batch), C.POINTER(Options), C.POINTER(BatchResult)]
    lib.xreac_evaluate_batch.restype = C.c_int
    lib.xreac_batch_error.argtypes = [vp, C.c_int32]
    lib.xreac_batch_error.restype = C.c_char_p
    lib.xreac_verlet.argtypes = [vp, vp, C.c_int32, vp, vp, vp, vp, vp, vp, C.c_int32, vp]
    lib.xreac_verlet.restype = C.c_int
    lib.xreac_fire_step.argtypes = [vp, C.c_int32, vp, vp, vp, vp, vp, vp, C.c_double, C.c_int32]
    lib.xreac_fire_step.restype = C.c_int
    lib.xreac_qeq_extrapolate.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, vp]
    lib.xreac_qeq_extrapolate.restype = C.c_int
    lib.xreac_neighbor_list_builds.argtypes = [vp]
    lib.xreac_neighbor_list_builds.restype = C.c_int64
    lib.xreac_copy_neighbors.argtypes = [vp, vp, vp, vp, C.c_int64]
    lib.xreac_copy_neighbors.restype = C.c_int
    lib.xreac_copy_bonds.argtypes = [vp, vp, vp, vp, vp, C.c_int64]
    lib.xreac_copy_bonds.restype = C.c_int
    lib.xreac_profile_enable.argtypes = [vp, C.c_int]
    lib.xreac_profile_enable.restype = C.c_int
    lib.xreac_profile_stride.argtypes = [vp, C.c_int]
    lib.xreac_profile_stride.restype = C.c_int
    lib.xreac_profile_read.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int]
    lib.xreac_profile_read.restype = C.c_int
    lib.xreac_step_create.argtypes = [vp, vp, C.POINTER(Batch), C.POINTER(Options), C.POINTER(StepStages),
                                      C.POINTER(BatchResult), C.POINTER(vp)]
    lib.xreac_step_create.restype = C.c_int
    lib.xreac_step_run.argtypes = [vp, C.POINTER(BatchResult)]
    lib.xreac_step_run.restype = C.c_int
    lib.xreac_step_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.xreac_step_stats.restype = C.c_int
    lib.xreac_step_destroy.argtypes = [vp]
    lib.xreac_step_destroy.restype = None
    _LIB = lib
    return lib
