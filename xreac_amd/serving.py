"""Continuous-batching request server on one GPU.

A request is one energy + forces + QEq evaluation, an NVE segment of
``md_steps`` velocity-Verlet steps (``md_steps + 1`` evaluations), or a FIRE
relaxation (``relax=True``, the reference's native loop,
src/xreac/calculator.py:213-246) that ends at the server's force tolerance or
iteration limit.  Requests
wait in a FIFO queue and are admitted at the start of every step while the
estimated device footprint of the active set stays within the budget (a
request larger than the budget is admitted alone).  Each step evaluates every
active request once, one ragged batch per force field, and a request leaves
the batch as soon as its work is done, so requests that arrive while others
are on the GPU join the next step.  Positions, velocities, forces and the QEq
(s, t) state of every active request stay on the device; the QEq solve of a
request starts from its previous solution.

The reference evaluates one system per call and has no dynamics
(src/xreac/calculator.py:76-165, docs/index.md:53-56); this is the serving
form of that call.
"""

from __future__ import annotations

import threading
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .engine import Engine
from .forcefield import ForceField
from .workloads import maxwell_boltzmann

BYTES_PER_ATOM = 4096  # per-atom workspace (CG vectors, bond-order state, outputs)
BYTES_PER_EDGE = 40  # neighbor CSR, H (FP64 + FP32), flags, bond slots
EDGES_PER_ATOM = 450  # liquid density within 10 A


def footprint(atoms, periodic):
    """Estimated device bytes of one request."""
    edges = atoms * (EDGES_PER_ATOM if periodic else min(max(atoms - 1, 1), EDGES_PER_ATOM))
    return atoms * BYTES_PER_ATOM + edges * BYTES_PER_EDGE


@dataclass
class Response:
    rid: int
    atoms: int
    evaluations: int
    energy: float
    components: np.ndarray
    forces: np.ndarray
    charges: np.ndarray
    arrival: float
    finished: float
    error: str | None = None
    positions: np.ndarray | None = None  # NVE segments: final state
    velocities: np.ndarray | None = None
    potential: np.ndarray | None = None  # per evaluation
    kinetic: np.ndarray | None = None
    converged: bool | None = None  # relaxations

    @property
    def latency(self):
        return self.finished - self.arrival


@dataclass
class _Job:
    req: dict
    n: int
    periodic: bool
    bytes: int
    arrival: float
    evals: int = 0
    potential: list = field(default_factory=list)
    kinetic: list = field(default_factory=list)
    prep: tuple | None = None  # host arrays for _Group.append, built ahead of admission
    admitted: bool = False


class _Group:
    """Device state of the active requests of one force field, packed in job order.

    Positions, velocities, previous forces, QEq state and FIRE state live in
    one set of tensors, so a step costs a fixed number of tensor operations
    whatever the number of requests: arrivals are appended with one copy per
    array, finished requests dropped with one gather.
    """

    def __init__(self, torch, eng):
        self.torch, self.eng, self.dev = torch, eng, eng.device
        self.jobs = []
        self.sizes = np.zeros(0, dtype=np.int64)
        self.md = np.zeros(0, dtype=bool)
        self.relax = np.zeros(0, dtype=bool)
        self.dt = np.zeros(0, dtype=np.float64)
        self.cells = np.zeros((0, 3, 3), dtype=np.float64)
        self.pbcs = np.zeros((0, 3), dtype=np.int32)
        self.charges = np.zeros(0, dtype=np.float64)
        f64 = dict(dtype=torch.float64, device=self.dev)
        self.X = torch.zeros((0, 3), **f64)
        self.V = torch.zeros((0, 3), **f64)
        self.F = torch.zeros((0, 3), **f64)
        self.QS = torch.zeros((0, 2), **f64)
        self.QH = torch.zeros((3, 0, 2), **f64)  # previous QEq solutions, newest first (MD extrapolation)
        self.T = torch.zeros(0, dtype=torch.int32, device=self.dev)
        self.FS = torch.zeros((0, 4), **f64)
        self.ST = torch.zeros(0, dtype=torch.int32, device=self.dev)
        self._mass = {}  # per force field: element -> mass
        # Composition: bumped by every append and drop.  A step whose group has
        # the composition of its previous step runs as a recorded step (one
        # graph replay), which needs fixed per-step buffers.
        self.version = 0
        self.stepped_version = -1
        self.rec = None
        self.Q = self.K = self.DT = self.ORD = None
        self._off, self._off_version = None, -1

    def prepare(self, job, ff):
        """Host arrays of one job (positions, type ids, velocities, FIRE state); pure host work."""
        req = job.req
        x = np.asarray(req["positions"], dtype=np.float64).reshape(job.n, 3)
        t = self.eng.type_ids(req["symbols"])
        if req.get("md_steps", 0) > 0:
            v0 = req.get("velocities")
            if v0 is None:
                mass = self._mass.get(id(ff))
                if mass is None:
                    mass = self._mass[id(ff)] = {s: a["mass"] for s, a in ff.atoms.items()}
                masses = np.fromiter(map(mass.__getitem__, req["symbols"]), dtype=np.float64, count=job.n)
                v0 = maxwell_boltzmann(masses, req.get("temperature", 300.0), req.get("velocity_seed", 0))
            v = np.asarray(v0, dtype=np.float64).reshape(job.n, 3)
        else:
            v = np.zeros((job.n, 3))
        relax = bool(req.get("relax"))
        fs = [0.02, 0.1, 0.0, 0.0] if relax else [0.0, 0.0, 0.0, 0.0]
        st = 0 if relax else 1  # status 1: not relaxing, FIRE leaves the system alone
        job.prep = (x, t, v, fs, st)

    def append(self, jobs, ff):
        """Pack newly admitted jobs behind the active ones (one host-to-device copy per array)."""
        torch = self.torch
        x, t, v, fs, st = [], [], [], [], []
        for job in jobs:
            if job.prep is None:
                self.prepare(job, ff)
            xj, tj, vj, fsj, stj = job.prep
            job.prep = None
            x.append(xj)
            t.append(tj)
            v.append(vj)
            fs.append(fsj)
            st.append(stj)
        m = sum(job.n for job in jobs)
        f64 = dict(dtype=torch.float64, device=self.dev)
        self.X = torch.cat([self.X, torch.as_tensor(np.concatenate(x), **f64)])
        self.V = torch.cat([self.V, torch.as_tensor(np.concatenate(v), **f64)])
        self.T = torch.cat([self.T, torch.as_tensor(np.concatenate(t), dtype=torch.int32, device=self.dev)])
        self.F = torch.cat([self.F, torch.zeros((m, 3), **f64)])
        self.QS = torch.cat([self.QS, torch.zeros((m, 2), **f64)])
        self.QH = torch.cat([self.QH, torch.zeros((3, m, 2), **f64)], dim=1)
        self.FS = torch.cat([self.FS, torch.as_tensor(np.array(fs, dtype=np.float64), **f64)])
        self.ST = torch.cat([self.ST, torch.as_tensor(np.array(st, dtype=np.int32), device=self.dev)])
        self.jobs.extend(jobs)
        self.sizes = np.concatenate([self.sizes, [job.n for job in jobs]])
        self.md = np.concatenate([self.md, [job.req.get("md_steps", 0) > 0 for job in jobs]])
        self.relax = np.concatenate([self.relax, [bool(job.req.get("relax")) for job in jobs]])
        self.dt = np.concatenate([self.dt, [job.req.get("dt", 0.25) for job in jobs]])
        self.cells = np.concatenate([self.cells, [np.eye(3) if job.req.get("cell") is None else
                                                  np.asarray(job.req["cell"], dtype=np.float64).reshape(3, 3)
                                                  for job in jobs]])
        self.pbcs = np.concatenate([self.pbcs, np.array([[0, 0, 0] if job.req.get("cell") is None else
                                                         [int(bool(p)) for p in job.req["pbc"]] for job in jobs],
                                                        dtype=np.int32).reshape(-1, 3)])
        self.charges = np.concatenate([self.charges, [job.req.get("total_charge", 0.0) for job in jobs]])
        self.version += 1

    def drop(self, done):
        """Remove the jobs flagged in ``done`` (bool per job) with one gather per array."""
        torch = self.torch
        keep = ~done
        atom_keep = np.repeat(keep, self.sizes)
        ia = torch.as_tensor(np.flatnonzero(atom_keep), device=self.dev)
        isys = torch.as_tensor(np.flatnonzero(keep), device=self.dev)
        self.X = self.X.index_select(0, ia)
        self.V = self.V.index_select(0, ia)
        self.F = self.F.index_select(0, ia)
        self.QS = self.QS.index_select(0, ia)
        self.QH = self.QH.index_select(1, ia)
        self.T = self.T.index_select(0, ia)
        self.FS = self.FS.index_select(0, isys)
        self.ST = self.ST.index_select(0, isys)
        self.jobs = [j for j, k in zip(self.jobs, keep) if k]
        for name in ("sizes", "md", "relax", "dt", "cells", "pbcs", "charges"):
            setattr(self, name, getattr(self, name)[keep])
        self.version += 1

    def offsets_dev(self, offsets):
        """Device copy of the packed offsets, made once per composition."""
        if self._off_version != self.version:
            self._off = self.torch.as_tensor(offsets, device=self.dev)
            self._off_version = self.version
        return self._off

    def forget_step(self):
        if self.rec is not None:
            self.rec.close()
            self.rec = None


class Server:
    """Continuous batching over one GPU, one engine per bundled force field.

    precision: "fp64" or "mixed" QEq; warm_start: start each QEq solve from the
    request's previous (s, t); extrapolate: for NVE segments, start instead
    from the polynomial through the last (up to four) solutions, one step
    ahead, which roughly halves the warm solve's iterations; neighbor_skin: keep
    each group's device-built neighbor list with this per-atom allowance (the
    reference ASE calculator's neighbor_skin) and reuse it while the group's
    batch keeps its shape and no atom has moved further (results bitwise those
    of fresh lists; pays for NVE-heavy work, while a batch that changes at
    every step rebuilds every step); recorded_steps: a group whose requests are
    the same as at its previous step runs the step as a recorded HIP graph
    (``Engine.record_step``; results bitwise those of the unrecorded step; off
    by default: replays measured slower than host-issued steps);
    max_systems: cap on requests per batch (1 gives
    the one-request-at-a-time plain path); budget_bytes / max_batch_atoms bound
    the active set; full_derivative: charge-response forces (the exact gradient
    of the QEq-equilibrated energy, calculator.py:79-86) for every request.
    """

    INLINE_PREPARE = 32  # unprepared requests a loop admits (and prepares itself) per step

    def __init__(self, force_fields=("ffield.reax.HO.2015", "ffield.reax.CHO.2008"), device=0, *,
                 precision="fp64", warm_start=True, extrapolate=True, neighbor_skin=0.0, tol=1e-10,
                 max_systems=None, max_batch_atoms=262144, budget_bytes=32 << 30, relax_tolerance=1e-4,
                 relax_iterations=500, full_derivative=False, recorded_steps=False, clock=time.perf_counter):
        import torch

        self.torch = torch
        self.engines, self._ff = {}, {}
        for name in force_fields:
            ff = ForceField.bundled(name)
            self._ff[name] = ff
            # one stream per force field (default priority): the groups' steps overlap on the device
            self.engines[name] = Engine(ff, device, stream=torch.cuda.Stream(device))
        self._pool = ThreadPoolExecutor(max_workers=max(1, len(self.engines)))
        self.precision = precision
        self.warm_start = warm_start
        self.extrapolate = bool(extrapolate)
        self.neighbor_skin = float(neighbor_skin)
        self.tol = tol
        self.max_systems = max_systems
        self.max_batch_atoms = max_batch_atoms
        self.budget = budget_bytes
        self.relax_tolerance = relax_tolerance
        self.relax_iterations = relax_iterations
        self.full_derivative = bool(full_derivative)
        self.recorded_steps = bool(recorded_steps)
        self.recorded_runs = 0  # group steps run as recorded steps (graph replays after the first)
        self.clock = clock
        self.queues = {k: deque() for k in self.engines}  # per force field, in arrival order
        self.active = {k: _Group(torch, eng) for k, eng in self.engines.items()}
        self._lock = threading.Lock()  # queues, budget and results, shared by the groups' loops
        self._closed = False
        self._unprepared = deque()  # submitted requests whose host arrays are not built yet, in arrival order
        self.used_bytes = 0
        self.responses = []
        self.atom_evals = 0
        self.steps = 0

    # ---------------------------------------------------------------- intake
    def submit(self, req, arrival=None):
        if req["force_field"] not in self.engines:
            raise ValueError(f"no engine for force field {req['force_field']!r}")
        n = len(req["symbols"])
        periodic = req.get("cell") is not None and any(req.get("pbc", (True, True, True)))
        job = _Job(req, n, periodic, footprint(n, periodic), self.clock() if arrival is None else arrival)
        with self._lock:
            self.queues[req["force_field"]].append(job)
            self._unprepared.append(job)

    def _prepare_ahead(self):
        """Build submitted requests' host arrays in arrival order while the loops run, so admitting many
        requests at once does not hold a loop (and the device) for their preparation (about 45 us each);
        a request admitted before its turn here is prepared by the loop itself."""
        while True:
            with self._lock:
                while self._unprepared and self._unprepared[0].admitted:
                    self._unprepared.popleft()
                job = self._unprepared.popleft() if self._unprepared else None
                closed = self._closed
            if job is None:
                if closed:
                    return
                time.sleep(2e-4)
                continue
            if job.prep is None:
                name = job.req["force_field"]
                self.active[name].prepare(job, self._ff[name])

    @property
    def queue(self):
        """Requests waiting for admission, all force fields."""
        return [job for q in self.queues.values() for job in q]

    def _admit(self, name):
        """Admit queued requests of force field ``name`` in arrival order while the shared budget allows
        (the caller holds the lock); a group with nothing active always takes its next request.  At most
        INLINE_PREPARE requests whose host arrays are not built yet are admitted per step (the loop builds
        them before its step): behind a burst of arrivals the device starts on a small batch at once and
        the batches grow as the preparer thread gets ahead, instead of the device idling while the loop
        prepares the whole burst."""
        group = self.active[name]
        queue = self.queues[name]
        active = sum(len(g.jobs) for g in self.active.values())
        atoms = sum(int(g.sizes.sum()) for g in self.active.values())
        admitted = []
        inline = 0
        while queue:
            job = queue[0]
            if self.max_systems is not None and active >= self.max_systems:
                break
            if (group.jobs or admitted) and (self.used_bytes + job.bytes > self.budget or
                                              atoms + job.n > self.max_batch_atoms):
                break
            if job.prep is None:
                if inline >= self.INLINE_PREPARE and (group.jobs or admitted):
                    break
                inline += 1
            queue.popleft()
            job.admitted = True
            admitted.append(job)
            self.used_bytes += job.bytes
            active += 1
            atoms += job.n
        return admitted

    def _loop(self, name):
        """One force field's serving loop: admit, evaluate its batch once, hand back finished requests.

        The groups' loops run side by side on their own streams, each at its
        own pace (the library releases the interpreter lock while it works),
        so one group never waits at the other's step boundary.
        """
        group = self.active[name]
        while True:
            with self._lock:
                admitted = self._admit(name)
                closed = self._closed and not self.queues[name]
            if admitted:
                with self.torch.cuda.stream(group.eng.stream):  # the group's tensors live on its stream
                    group.append(admitted, self._ff[name])
            if not group.jobs:
                if closed:
                    return
                time.sleep(2e-4)
                continue
            evals, responses, freed = self._run_group(group)
            with self._lock:
                self.atom_evals += evals
                self.responses.extend(responses)
                self.used_bytes -= freed
                self.steps += 1

    def _run_group(self, g):
        """One step of a group on its engine's stream; returns (atom evaluations, finished responses, bytes freed)."""
        with self.torch.cuda.stream(g.eng.stream):
            return self._step_group(g)

    def _step_group(self, g):
        torch = self.torch
        eng, dev = g.eng, g.dev
        ns = len(g.jobs)
        offsets = np.zeros(ns + 1, dtype=np.int32)
        np.cumsum(g.sizes, out=offsets[1:])
        n = int(offsets[-1])
        evals = np.fromiter((j.evals for j in g.jobs), dtype=np.int64, count=ns)
        moving = g.md & (evals > 0)
        md_any, relax_any = bool(g.md.any()), bool(g.relax.any())
        extrap = self.warm_start and self.extrapolate
        order = np.where(moving, np.minimum(evals, 4) - 1, 0).astype(np.int32)
        dt = np.where(moving, g.dt, 0.0)
        stable = g.version == g.stepped_version
        g.stepped_version = g.version
        if not stable:  # a new composition: fresh per-step buffers, no recorded step
            g.forget_step()
            f64 = dict(dtype=torch.float64, device=dev)
            g.Q = torch.empty(n, **f64)
            g.K = torch.zeros(ns, **f64)
            g.DT = torch.zeros(ns, **f64)
            g.ORD = torch.zeros(ns, dtype=torch.int32, device=dev)
        g.DT.copy_(torch.from_numpy(dt))
        g.ORD.copy_(torch.from_numpy(order))
        out = dict(forces=g.F, charges=g.Q, total_bond_orders=None, lone_pairs=None, bond_counts=None,
                   qeq_state=g.QS)
        # The same composition as the last step: the step runs recorded (its
        # first run records it, later ones replay it).  With every MD request
        # past its first evaluation the recorded verlet and extrapolation
        # stages are exactly the ones the unrecorded step below issues.
        recorded = (stable and self.recorded_steps and self.neighbor_skin == 0.0 and
                    (not md_any or bool(moving[g.md].all())))
        if recorded:
            if g.rec is None:
                g.rec = eng.record_step(
                    g.X, g.T, offsets, g.cells, g.pbcs, g.charges, out=out, tol=self.tol, precision=self.precision,
                    warm_start=self.warm_start, full_derivative=self.full_derivative,
                    dt_fs=g.DT if md_any else None, velocities=g.V if (md_any or relax_any) else None,
                    kinetic=g.K if md_any else None, qeq_order=g.ORD if (md_any and extrap) else None,
                    qeq_history=g.QH if (md_any and extrap) else None, fire_state=g.FS if relax_any else None,
                    fire_status=g.ST if relax_any else None, fire_force_tolerance=self.relax_tolerance,
                    fire_max_iterations=self.relax_iterations)
            res = g.rec.run()
            self.recorded_runs += 1
        else:
            g.forget_step()
            if moving.any():
                eng.verlet(g.offsets_dev(offsets), g.T, g.DT, g.X, g.V, g.F, 0)
                if extrap:
                    # degree min(evaluations, 4) - 1 through the segment's solutions so far (fixed dt)
                    eng.qeq_extrapolate(g.offsets_dev(offsets), g.ORD, g.QS, g.QH)
            res = eng.evaluate_batch(g.X, g.T, offsets, g.cells, g.pbcs, g.charges, tol=self.tol,
                                     precision=self.precision, warm_start=self.warm_start,
                                     full_derivative=self.full_derivative, out=out, neighbor_skin=self.neighbor_skin)
            if md_any:
                eng.verlet(g.offsets_dev(offsets), g.T, g.DT, g.X, g.V, g.F, 1, g.K)
            if relax_any:
                eng.fire_step(g.offsets_dev(offsets), g.X, g.V, g.F, g.FS, g.ST, self.relax_tolerance,
                              self.relax_iterations)
        kin = g.K.cpu().numpy() if md_any else None
        fire_status = g.ST.cpu().numpy() if relax_any else None
        comps = res["components"]
        pot = comps.sum(axis=1)
        errors = res["errors"]
        done = np.zeros(ns, dtype=bool)
        for k, j in enumerate(g.jobs):
            j.evals += 1
            j.potential.append(float(pot[k]))
            if g.md[k]:
                j.kinetic.append(float(kin[k]))
            finished = fire_status[k] != 0 if g.relax[k] else j.evals > j.req.get("md_steps", 0)
            done[k] = errors[k] is not None or finished
        if not done.any():
            return n, [], 0
        F = out["forces"].cpu().numpy()
        Q = out["charges"].cpu().numpy()
        Xh = g.X.cpu().numpy() if (g.md | g.relax)[done].any() else None
        Vh = g.V.cpu().numpy() if g.md[done].any() else None
        now = self.clock()
        finished, freed = [], 0
        for k in np.flatnonzero(done):
            j = g.jobs[k]
            lo, hi = offsets[k], offsets[k + 1]
            r = Response(j.req["rid"], j.n, j.evals, float(pot[k]), comps[k].copy(), F[lo:hi].copy(),
                         Q[lo:hi].copy(), j.arrival, now, errors[k])
            if g.md[k]:
                r.positions = Xh[lo:hi].copy()
                r.velocities = Vh[lo:hi].copy()
                r.potential = np.array(j.potential)
                r.kinetic = np.array(j.kinetic)
            if g.relax[k]:
                # FIRE leaves a converged (or exhausted) system where it was evaluated
                r.positions = Xh[lo:hi].copy()
                r.potential = np.array(j.potential)
                r.converged = bool(fire_status[k] == 1)
            finished.append(r)
            freed += j.bytes
        g.drop(done)
        return n, finished, freed

    # ------------------------------------------------------------------- run
    def idle(self):
        return not any(self.queues.values()) and not any(g.jobs for g in self.active.values())

    def run(self, requests, arrivals=None):
        """Serve ``requests`` arriving at ``arrivals`` (seconds from now; all at once if None).

        Returns the responses in completion order; ``steps`` counts the batched
        evaluations of all groups.
        """
        t0 = self.clock()
        order = np.argsort(arrivals, kind="stable") if arrivals is not None else np.arange(len(requests))
        times = np.zeros(len(requests)) if arrivals is None else np.asarray(arrivals, dtype=float)
        with self._lock:
            self._closed = False
        preparer = threading.Thread(target=self._prepare_ahead, daemon=True)
        preparer.start()
        loops = [self._pool.submit(self._loop, name) for name in self.active]
        i = 0
        try:
            while i < len(order):
                now = self.clock() - t0
                while i < len(order) and times[order[i]] <= now:
                    self.submit(requests[order[i]], arrival=t0 + times[order[i]])
                    i += 1
                if i < len(order):
                    time.sleep(min(max(times[order[i]] - (self.clock() - t0), 0.0), 0.001))
        finally:
            with self._lock:
                self._closed = True
            for f in loops:
                f.result()
            preparer.join()
        return self.responses
