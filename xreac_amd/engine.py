"""Device engine: one force field bound to one GPU context and stream.

Thin owner of the C handles of ``libxreac_hip.so``.  Inputs and outputs are
device tensors; PyTorch only supplies device memory and the stream.
"""

from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _lib
from .tables import build_tables

COMPONENTS = (
    "bond",
    "atom",
    "lone_pair",
    "molecule",
    "angle",
    "penalty",
    "angle_conjugation",
    "hydrogen_bond",
    "torsion",
    "conjugation",
    "vdw",
    "coulomb",
    "electric_field",
    "qeq",
)


def _ptr(a):
    return C.c_void_p(a.ctypes.data if isinstance(a, np.ndarray) else int(a.data_ptr()))


def _precision(name):
    try:
        return {"fp64": _lib.PRECISION_FP64, "mixed": _lib.PRECISION_MIXED}[name]
    except KeyError:
        raise ValueError(f"precision must be 'fp64' or 'mixed', not {name!r}") from None


class DeviceError(RuntimeError):
    pass


class Engine:
    def __init__(self, force_field, device=0, stream=None):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("xreac_amd evaluates on an AMD GPU; no ROCm device is visible")
        self.torch = torch
        self.lib = _lib.load()
        self.device = torch.device("cuda", device)
        self.tables = build_tables(force_field)
        t = self.tables
        keep = [
            t.general, t.atom, t.pair, t.angle_offsets, t.angle_rows if len(t.angle_rows) else np.zeros((1, 7)),
            t.torsion, t.torsion_on, t.hbond, t.hbond_on,
        ]
        tab = _lib.FFTables(
            len(t.labels), t.vdw_type, *[C.c_void_p(a.ctypes.data) for a in keep]
        )
        self._ff = C.c_void_p()
        self._check(self.lib.xreac_ff_create(C.byref(tab), C.byref(self._ff)), None)
        # the context launches on `stream` (a torch.cuda.Stream; default: the current stream)
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        self._ctx = C.c_void_p()
        self._steps = weakref.WeakSet()  # recorded steps, closed before the context
        with torch.cuda.device(self.device):
            self._check(self.lib.xreac_ctx_create(device, C.c_void_p(self.stream.cuda_stream), C.byref(self._ctx)),
                        None)

    def close(self):
        for step in list(getattr(self, "_steps", ())):
            step.close()
        if getattr(self, "_ctx", None):
            self.lib.xreac_ctx_destroy(self._ctx)
            self._ctx = None
        if getattr(self, "_ff", None):
            self.lib.xreac_ff_destroy(self._ff)
            self._ff = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, status, ctx):
        if status == _lib.XREAC_OK:
            return
        msg = self.lib.xreac_error_message(ctx).decode() if ctx is not None else "xreac call failed"
        if status == _lib.XREAC_INVALID:
            raise ValueError(msg)
        raise DeviceError(msg)

    def type_ids(self, symbols):
        return np.fromiter(map(self.tables.index.__getitem__, symbols), dtype=np.int32, count=len(symbols))

    def outputs(self, n, properties=True):
        """Device output buffers for an n-atom evaluation (reusable across calls via ``out=``)."""
        torch = self.torch
        dev = self.device
        out = dict(
            forces=torch.empty((n, 3), dtype=torch.float64, device=dev),
            charges=torch.empty(n, dtype=torch.float64, device=dev),
            total_bond_orders=None,
            lone_pairs=None,
            bond_counts=None,
        )
        if properties:
            out["total_bond_orders"] = torch.empty(n, dtype=torch.float64, device=dev)
            out["lone_pairs"] = torch.empty(n, dtype=torch.float64, device=dev)
            out["bond_counts"] = torch.empty(n, dtype=torch.int32, device=dev)
        return out

    def evaluate(self, positions, types, cell=None, pbc=(False, False, False), total_charge=0.0, *, tol=1e-10,
                 max_iter=2000, lammps_angles=True, properties=True, out=None, precision="fp64",
                 full_derivative=False, charges=None, neighbors=None, no_alloc=False, neighbor_skin=0.0):
        """Evaluate one system whose positions (n,3 float64) and types (n int32) are device tensors.

        Returns a dict of device tensors (forces, charges and optionally
        total_bond_orders, lone_pairs, bond_counts), the 14 components (host
        numpy) and the edge/bond/iteration counts.  ``out`` (from
        ``outputs``) is filled in place instead of allocating new buffers.
        ``neighbors`` = canonical CSR ``(row_ptr, j, S)`` (host or device
        int32, see ``geometry.canonical_neighbors``) replaces the device build.
        ``no_alloc`` runs ``xreac_evaluate_static``: the library takes every
        workspace from its static device scratch and allocates nothing (pass
        ``out`` to keep the call itself free of PyTorch allocations too).
        ``neighbor_skin`` > 0 keeps the device-built list (within the cutoff
        plus twice the skin) on this engine and reuses it on later calls of
        the same shape until an atom has moved more than the skin; results
        are bitwise those of a fresh list (``neighbor_list_builds`` counts
        the builds).
        """
        torch = self.torch
        n = int(positions.shape[0])
        assert positions.dtype == torch.float64 and positions.is_contiguous() and positions.shape == (n, 3)
        assert types.dtype == torch.int32 and types.is_contiguous() and types.shape == (n,)
        if out is None:
            out = self.outputs(n, properties)
        given = charges
        forces, charges = out["forces"], out["charges"]
        if given is not None:
            # evaluate at these charges instead of solving QEq
            charges.copy_(self.torch.as_tensor(given, dtype=self.torch.float64, device=self.device).reshape(n))
        tbo, lp, bc = out["total_bond_orders"], out["lone_pairs"], out["bond_counts"]
        assert forces.shape == (n, 3) and charges.shape == (n,)
        cell_arr = None
        if cell is not None:
            cell_arr = np.ascontiguousarray(cell, dtype=np.float64).reshape(3, 3)
        nbr = self._device_csr(neighbors, n) if neighbors is not None else None
        sysd = _lib.System(
            n,
            _ptr(positions),
            _ptr(types),
            C.c_void_p(cell_arr.ctypes.data) if cell_arr is not None else None,
            (C.c_int32 * 3)(*[int(bool(v)) for v in pbc]),
            float(total_charge),
            *(self._csr_fields(nbr)),
        )
        opt = _lib.Options(float(tol), int(max_iter), 1 if lammps_angles else 0, _precision(precision), 0,
                           1 if full_derivative else 0, 1 if given is not None else 0, float(neighbor_skin))
        res = _lib.Result()
        res.forces = _ptr(forces)
        res.charges = _ptr(charges)
        res.total_bond_orders = _ptr(tbo) if tbo is not None else None
        res.lone_pairs = _ptr(lp) if lp is not None else None
        res.bond_counts = _ptr(bc) if bc is not None else None
        entry = self.lib.xreac_evaluate_static if no_alloc else self.lib.xreac_evaluate
        status = entry(self._ctx, self._ff, C.byref(sysd), C.byref(opt), C.byref(res))
        self._check(status, self._ctx)
        return dict(
            forces=forces,
            charges=charges,
            total_bond_orders=tbo,
            lone_pairs=lp,
            bond_counts=bc,
            components=np.array(res.components[:], dtype=np.float64),
            num_edges=int(res.num_edges),
            num_bonds=int(res.num_bonds),
            qeq_iterations=int(res.qeq_iterations),
            n=n,
        )

    def _device_csr(self, neighbors, n):
        """Device int32 copies of a canonical CSR (row_ptr [n+1], j [M], S [M, 3]).

        The CSR is checked on the host before any kernel runs: row_ptr starts
        at 0, never decreases and ends at len(j); every j is an atom index; every
        shift fits the device's 10-bit fields.  (The device clamps the pointers
        again, so a list passed straight to the C ABI cannot reach outside its
        buffers either.)
        """
        torch = self.torch
        host = [np.asarray(a.cpu() if hasattr(a, "cpu") else a) for a in neighbors]
        if len(host) != 3:
            raise ValueError("neighbors must be a CSR (row_ptr [n + 1], j [M], S [M, 3])")
        row_ptr, j, S = host[0].reshape(-1), host[1].reshape(-1), host[2].reshape(-1, 3)
        if row_ptr.shape[0] != n + 1 or S.shape[0] != j.shape[0]:
            raise ValueError("neighbors must be a CSR (row_ptr [n + 1], j [M], S [M, 3])")
        if row_ptr[0] != 0 or np.any(np.diff(row_ptr.astype(np.int64)) < 0) or row_ptr[-1] != len(j):
            raise ValueError("neighbor row pointers must start at 0, never decrease and end at the number of edges")
        if len(j) and (j.min() < 0 or j.max() >= n):
            raise ValueError("Neighbor atom indices are out of range")
        if len(j) and np.max(np.abs(S.astype(np.int64))) > 511:
            raise ValueError("Neighbor shifts beyond 511 cells are not supported on the device")
        return [torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=self.device) for a in (row_ptr, j, S)]

    @staticmethod
    def _csr_fields(nbr):
        if nbr is None:
            return None, None, None, 0
        return _ptr(nbr[0]), _ptr(nbr[1]), _ptr(nbr[2]), int(nbr[1].shape[0])

    def batch_outputs(self, n, properties=False, qeq_state=False):
        """Device output buffers for a batch of n atoms in total (reusable via ``out=``)."""
        out = self.outputs(n, properties)
        out["qeq_state"] = (
            self.torch.zeros((n, 2), dtype=self.torch.float64, device=self.device) if qeq_state else None
        )
        return out

    def evaluate_batch(self, positions, types, offsets, cells=None, pbc=None, total_charges=None, *, tol=1e-10,
                       max_iter=2000, lammps_angles=True, properties=False, precision="fp64", warm_start=False,
                       full_derivative=False, out=None, neighbors=None, neighbor_skin=0.0):
        """Evaluate independent systems packed atom-wise in one pipeline.

        positions (N, 3) float64 and types (N,) int32 are device tensors;
        ``offsets`` (ns + 1,) host integers delimit the systems; ``cells``
        (ns, 3, 3) and ``pbc`` (ns, 3) are host arrays (None: all
        nonperiodic); ``total_charges`` (ns,) or None.  ``out`` may carry a
        (N, 2) ``qeq_state`` tensor: with ``warm_start`` it is the CG starting
        guess, and it always receives the (s, t) solution.

        Returns device forces/charges (and properties), host ``components``
        (ns, 14), ``status`` (ns,), ``errors`` (message or None per system),
        ``qeq_iterations`` (ns,) and edge/bond counts.  A failed system does
        not affect the others.
        """
        torch = self.torch
        offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        ns = len(offsets) - 1
        n = int(offsets[-1])
        assert ns >= 1 and positions.shape == (n, 3) and types.shape == (n,)
        assert positions.dtype == torch.float64 and positions.is_contiguous()
        assert types.dtype == torch.int32 and types.is_contiguous()
        if out is None:
            out = self.batch_outputs(n, properties, qeq_state=warm_start)
        qs = out.get("qeq_state")
        if warm_start and qs is None:
            raise ValueError("warm_start needs out['qeq_state']")
        keep = []
        cells_arr = pbc_arr = q_arr = None
        if cells is not None:
            cells_arr = np.ascontiguousarray(cells, dtype=np.float64).reshape(ns, 9)
            keep.append(cells_arr)
            pbc_arr = np.ones((ns, 3), dtype=np.int32) if pbc is None else np.ascontiguousarray(pbc, dtype=np.int32)
            keep.append(pbc_arr)
        if total_charges is not None:
            q_arr = np.ascontiguousarray(total_charges, dtype=np.float64).reshape(ns)
        nbr = self._device_csr(neighbors, n) if neighbors is not None else None
        batch = _lib.Batch(
            ns,
            _ptr(offsets),
            _ptr(positions),
            _ptr(types),
            _ptr(cells_arr) if cells_arr is not None else None,
            _ptr(pbc_arr) if pbc_arr is not None else None,
            _ptr(q_arr) if q_arr is not None else None,
            *(self._csr_fields(nbr)),
        )
        comps = np.zeros((ns, _lib.NCOMP), dtype=np.float64)
        status = np.zeros(ns, dtype=np.int32)
        iters = np.zeros(ns, dtype=np.int32)
        res = _lib.BatchResult()
        res.components = _ptr(comps)
        res.status = _ptr(status)
        res.qeq_iterations = _ptr(iters)
        res.forces = _ptr(out["forces"])
        res.charges = _ptr(out["charges"])
        for name in ("total_bond_orders", "lone_pairs", "bond_counts", "qeq_state"):
            t = out.get(name)
            setattr(res, name, _ptr(t) if t is not None else None)
        opt = _lib.Options(float(tol), int(max_iter), 1 if lammps_angles else 0, _precision(precision),
                           1 if warm_start else 0, 1 if full_derivative else 0, 0, float(neighbor_skin))
        rc = self.lib.xreac_evaluate_batch(self._ctx, self._ff, C.byref(batch), C.byref(opt), C.byref(res))
        self._check(rc, self._ctx)
        errors = [None if status[k] == _lib.XREAC_OK else self.lib.xreac_batch_error(self._ctx, k).decode()
                  for k in range(ns)]
        result = {k: out.get(k) for k in ("forces", "charges", "total_bond_orders", "lone_pairs", "bond_counts",
                                          "qeq_state")}
        result.update(components=comps, status=status, errors=errors, qeq_iterations=iters,
                      num_edges=int(res.num_edges), num_bonds=int(res.num_bonds), offsets=offsets)
        return result

    def record_step(self, positions, types, offsets, cells=None, pbc=None, total_charges=None, *, out, tol=1e-10,
                    max_iter=2000, lammps_angles=True, precision="fp64", warm_start=False, full_derivative=False,
                    dt_fs=None, velocities=None, kinetic=None, qeq_order=None, qeq_history=None, fire_state=None,
                    fire_status=None, fire_force_tolerance=0.0, fire_max_iterations=0):
        """One step of a fixed batch, recorded as a HIP graph on its first run and replayed after.

        The step runs, in order: velocity-Verlet stage 0 (``dt_fs`` given),
        QEq extrapolation (``qeq_order`` given), the batch evaluation as
        ``evaluate_batch`` into ``out``, velocity-Verlet stage 1, and one FIRE
        iteration (``fire_state`` given).  Every tensor is used in place by
        every run: contents may change between runs, shapes and storage not.
        The engine needs its own stream (``Engine(..., stream=torch.cuda.Stream())``).
        """
        torch = self.torch
        if self.stream.cuda_stream == 0:
            raise ValueError("record_step needs an Engine created with its own stream")
        offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        ns = len(offsets) - 1
        n = int(offsets[-1])
        assert positions.shape == (n, 3) and positions.dtype == torch.float64 and positions.is_contiguous()
        assert types.shape == (n,) and types.dtype == torch.int32 and types.is_contiguous()
        keep = [offsets, positions, types, out]
        cells_arr = pbc_arr = q_arr = None
        if cells is not None:
            cells_arr = np.ascontiguousarray(cells, dtype=np.float64).reshape(ns, 9)
            pbc_arr = np.ones((ns, 3), dtype=np.int32) if pbc is None else np.ascontiguousarray(pbc, dtype=np.int32)
        if total_charges is not None:
            q_arr = np.ascontiguousarray(total_charges, dtype=np.float64).reshape(ns)
        batch = _lib.Batch(ns, _ptr(offsets), _ptr(positions), _ptr(types),
                           _ptr(cells_arr) if cells_arr is not None else None,
                           _ptr(pbc_arr) if pbc_arr is not None else None,
                           _ptr(q_arr) if q_arr is not None else None, None, None, None, 0)
        res = _lib.BatchResult()
        res.forces = _ptr(out["forces"])
        res.charges = _ptr(out["charges"])
        for name in ("total_bond_orders", "lone_pairs", "bond_counts", "qeq_state"):
            t = out.get(name)
            setattr(res, name, _ptr(t) if t is not None else None)
        stages = _lib.StepStages()
        for name, t in (("dt_fs", dt_fs), ("velocities", velocities), ("kinetic", kinetic), ("qeq_order", qeq_order),
                        ("qeq_history", qeq_history), ("fire_state", fire_state), ("fire_status", fire_status)):
            if t is not None:
                assert t.is_cuda and t.is_contiguous(), name
                setattr(stages, name, _ptr(t))
                keep.append(t)
        stages.fire_force_tolerance = float(fire_force_tolerance)
        stages.fire_max_iterations = int(fire_max_iterations)
        opt = _lib.Options(float(tol), int(max_iter), 1 if lammps_angles else 0, _precision(precision),
                           1 if warm_start else 0, 1 if full_derivative else 0, 0, 0.0)
        handle = C.c_void_p()
        rc = self.lib.xreac_step_create(self._ctx, self._ff, C.byref(batch), C.byref(opt), C.byref(stages),
                                        C.byref(res), C.byref(handle))
        self._check(rc, self._ctx)
        step = RecordedStep(self, handle, ns, offsets, out, keep)
        self._steps.add(step)
        return step

    def verlet(self, offsets, types, dt_fs, positions, velocities, forces, stage, kinetic=None):
        """One velocity-Verlet stage on device tensors (offsets int32 [ns+1], dt_fs float64 [ns]).

        stage 0 kicks by dt/2 and drifts by dt; stage 1 kicks by dt/2 with
        the new forces and writes per-system kinetic energy (kcal/mol).
        """
        ns = int(offsets.shape[0]) - 1
        for t in (offsets, types, dt_fs, positions, velocities, forces):
            assert t.is_cuda and t.is_contiguous()
        rc = self.lib.xreac_verlet(self._ctx, self._ff, ns, _ptr(offsets), _ptr(types), _ptr(dt_fs),
                                   _ptr(positions), _ptr(velocities), _ptr(forces), int(stage),
                                   _ptr(kinetic) if kinetic is not None else None)
        self._check(rc, self._ctx)

    @property
    def neighbor_list_builds(self):
        """Skin lists built on this engine (one per system and rebuild), as ReaxFFCalculator counts them."""
        return int(self.lib.xreac_neighbor_list_builds(self._ctx))

    def qeq_extrapolate(self, offsets, order, qeq_state, history):
        """Extrapolated QEq starting guesses for the next equal MD step (device tensors: offsets int32
        [ns+1], order int32 [ns] in 0..3, qeq_state float64 [N, 2], history float64 [3, N, 2], newest
        first); qeq_state becomes the guess and history shifts by one."""
        ns = int(offsets.shape[0]) - 1
        n = int(qeq_state.shape[0])
        for t in (offsets, order, qeq_state, history):
            assert t.is_cuda and t.is_contiguous()
        assert order.shape == (ns,) and qeq_state.shape == (n, 2) and history.shape == (3, n, 2)
        rc = self.lib.xreac_qeq_extrapolate(self._ctx, ns, n, _ptr(offsets), _ptr(order), _ptr(qeq_state),
                                            _ptr(history))
        self._check(rc, self._ctx)

    def fire_step(self, offsets, positions, velocities, forces, state, status, force_tolerance, max_iterations):
        """One FIRE iteration per system (device tensors; state float64 [ns, 4], status int32 [ns])."""
        ns = int(offsets.shape[0]) - 1
        rc = self.lib.xreac_fire_step(self._ctx, ns, _ptr(offsets), _ptr(positions), _ptr(velocities),
                                      _ptr(forces), _ptr(state), _ptr(status), float(force_tolerance),
                                      int(max_iterations))
        self._check(rc, self._ctx)

    def profile(self, enable=True, timers=None, every=1):
        """Enable hipEvent timing of the named kernels (all when None) on one dispatch in ``every``, or disable it."""
        mask = 0
        if enable:
            names = _lib.TIMERS if timers is None else timers
            for name in names:
                mask |= 1 << _lib.TIMERS.index(name)
        self._check(self.lib.xreac_profile_stride(self._ctx, int(every)), self._ctx)
        self._check(self.lib.xreac_profile_enable(self._ctx, mask), self._ctx)

    def timers(self, reset=True):
        """{timer: (total_ms, launches)} accumulated since the last reset."""
        out = {}
        for k, name in enumerate(_lib.TIMERS):
            ms, cnt = C.c_double(), C.c_int64()
            self._check(self.lib.xreac_profile_read(self._ctx, k, C.byref(ms), C.byref(cnt), int(reset)), self._ctx)
            out[name] = (ms.value, cnt.value)
        return out

    def neighbors(self, n, m):
        """Canonical (row_ptr, j, S) of the last evaluation, host arrays."""
        row_ptr = np.zeros(n + 1, dtype=np.int32)
        j = np.zeros(max(m, 1), dtype=np.int32)
        S = np.zeros((max(m, 1), 3), dtype=np.int32)
        self._check(self.lib.xreac_copy_neighbors(self._ctx, _ptr(row_ptr), _ptr(j), _ptr(S), max(m, 1)), self._ctx)
        return row_ptr, j[:m], S[:m]

    def bonds(self, nb):
        """Active directed bonds (i, j, S, BO) of the last evaluation, host arrays."""
        i = np.zeros(max(nb, 1), dtype=np.int32)
        j = np.zeros(max(nb, 1), dtype=np.int32)
        S = np.zeros((max(nb, 1), 3), dtype=np.int32)
        bo = np.zeros(max(nb, 1), dtype=np.float64)
        self._check(
            self.lib.xreac_copy_bonds(self._ctx, _ptr(i), _ptr(j), _ptr(S), _ptr(bo), max(nb, 1)), self._ctx
        )
        return i[:nb], j[:nb], S[:nb], bo[:nb]


class RecordedStep:
    """A batch step recorded as a HIP graph (``Engine.record_step``)."""

    def __init__(self, engine, handle, ns, offsets, out, keep):
        self.engine, self._h, self.ns, self.offsets, self.out, self._keep = engine, handle, ns, offsets, out, keep
        self.components = np.zeros((ns, _lib.NCOMP), dtype=np.float64)
        self.status = np.zeros(ns, dtype=np.int32)
        self.iters = np.zeros(ns, dtype=np.int32)
        self._res = _lib.BatchResult()
        self._res.components = _ptr(self.components)
        self._res.status = _ptr(self.status)
        self._res.qeq_iterations = _ptr(self.iters)

    def run(self):
        """Run the step once (replayed after the first run) and wait; results as ``Engine.evaluate_batch``."""
        eng = self.engine
        rc = eng.lib.xreac_step_run(self._h, C.byref(self._res))
        eng._check(rc, eng._ctx)
        errors = [None if self.status[k] == _lib.XREAC_OK else eng.lib.xreac_batch_error(eng._ctx, k).decode()
                  for k in range(self.ns)]
        result = {k: self.out.get(k) for k in ("forces", "charges", "total_bond_orders", "lone_pairs", "bond_counts",
                                               "qeq_state")}
        result.update(components=self.components.copy(), status=self.status.copy(), errors=errors,
                      qeq_iterations=self.iters.copy(), num_edges=int(self._res.num_edges),
                      num_bonds=int(self._res.num_bonds), offsets=self.offsets)
        return result

    def stats(self):
        """(replays, fallbacks to an unrecorded run, recordings) so far."""
        r, f, c = C.c_int64(), C.c_int64(), C.c_int64()
        self.engine._check(self.engine.lib.xreac_step_stats(self._h, C.byref(r), C.byref(f), C.byref(c)), None)
        return r.value, f.value, c.value

    def close(self):
        if getattr(self, "_h", None):
            self.engine.lib.xreac_step_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
