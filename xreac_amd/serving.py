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

import time
from collections import deque
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
    x: object = None
    v: object = None
    types: object = None
    forces: object = None
    qstate: object = None
    fire: object = None  # relaxations: FIRE state (1, 4) and status (1,) on the device
    status: object = None
    evals: int = 0
    potential: list = field(default_factory=list)
    kinetic: list = field(default_factory=list)


class Server:
    """Continuous batching over one GPU, one engine per bundled force field.

    precision: "fp64" or "mixed" QEq; warm_start: start each QEq solve from the
    request's previous (s, t); max_systems: cap on requests per batch (1 gives
    the one-request-at-a-time plain path); budget_bytes / max_batch_atoms bound
    the active set.
    """

    def __init__(self, force_fields=("ffield.reax.HO.2015", "ffield.reax.CHO.2008"), device=0, *,
                 precision="fp64", warm_start=True, tol=1e-10, max_systems=None, max_batch_atoms=65536,
                 budget_bytes=32 << 30, relax_tolerance=1e-4, relax_iterations=500, clock=time.perf_counter):
        import torch

        self.torch = torch
        self.engines, self._ff = {}, {}
        for name in force_fields:
            ff = ForceField.bundled(name)
            self._ff[name] = ff
            self.engines[name] = Engine(ff, device)
        self.precision = precision
        self.warm_start = warm_start
        self.tol = tol
        self.max_systems = max_systems
        self.max_batch_atoms = max_batch_atoms
        self.budget = budget_bytes
        self.relax_tolerance = relax_tolerance
        self.relax_iterations = relax_iterations
        self.clock = clock
        self.queue = deque()
        self.active = {k: [] for k in self.engines}
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
        self.queue.append(job)

    def _admit(self):
        torch = self.torch
        active = sum(len(v) for v in self.active.values())
        atoms = sum(j.n for v in self.active.values() for j in v)
        while self.queue:
            job = self.queue[0]
            if self.max_systems is not None and active >= self.max_systems:
                break
            if active and (self.used_bytes + job.bytes > self.budget or atoms + job.n > self.max_batch_atoms):
                break
            self.queue.popleft()
            eng = self.engines[job.req["force_field"]]
            dev = eng.device
            job.x = torch.as_tensor(np.asarray(job.req["positions"], dtype=np.float64), device=dev).contiguous()
            job.types = torch.as_tensor(eng.type_ids(job.req["symbols"]), device=dev)
            if job.req.get("md_steps", 0) > 0:
                v0 = job.req.get("velocities")
                if v0 is None:
                    ff = self._ff[job.req["force_field"]]
                    masses = [ff.atoms[s]["mass"] for s in job.req["symbols"]]
                    v0 = maxwell_boltzmann(masses, job.req.get("temperature", 300.0), job.req.get("velocity_seed", 0))
                job.v = torch.as_tensor(np.asarray(v0, dtype=np.float64), device=dev).contiguous()
            elif job.req.get("relax"):
                job.v = torch.zeros((job.n, 3), dtype=torch.float64, device=dev)
                job.fire = torch.tensor([[0.02, 0.1, 0.0, 0.0]], dtype=torch.float64, device=dev)
                job.status = torch.zeros(1, dtype=torch.int32, device=dev)
            self.active[job.req["force_field"]].append(job)
            self.used_bytes += job.bytes
            active += 1
            atoms += job.n

    # ------------------------------------------------------------------ step
    def step(self):
        """Admit what fits, then evaluate every active request once."""
        self._admit()
        for key, jobs in self.active.items():
            if jobs:
                self._run_group(self.engines[key], jobs)
        self.steps += 1

    def _run_group(self, eng, jobs):
        torch = self.torch
        dev = eng.device
        ns = len(jobs)
        sizes = [j.n for j in jobs]
        offsets = np.zeros(ns + 1, dtype=np.int32)
        np.cumsum(sizes, out=offsets[1:])
        n = int(offsets[-1])
        X = torch.cat([j.x for j in jobs]).contiguous()
        T = torch.cat([j.types for j in jobs]).contiguous()
        zeros3 = lambda m: torch.zeros((m, 3), dtype=torch.float64, device=dev)  # noqa: E731
        md = [j.req.get("md_steps", 0) > 0 for j in jobs]
        relax = [bool(j.req.get("relax")) for j in jobs]
        moving = [m and j.evals > 0 for m, j in zip(md, jobs)]
        V = (torch.cat([j.v if j.v is not None else zeros3(j.n) for j in jobs]).contiguous()
             if any(md) or any(relax) else None)
        QS = torch.cat([j.qstate if j.qstate is not None else torch.zeros((j.n, 2), dtype=torch.float64, device=dev)
                        for j in jobs]).contiguous()
        cells = np.array([np.eye(3) if j.req.get("cell") is None else np.asarray(j.req["cell"], dtype=np.float64)
                          .reshape(3, 3) for j in jobs])
        pbcs = np.array([[0, 0, 0] if j.req.get("cell") is None else [int(bool(p)) for p in j.req["pbc"]]
                         for j in jobs], dtype=np.int32)
        charges = np.array([j.req.get("total_charge", 0.0) for j in jobs], dtype=np.float64)
        off_dev = dt_dev = kin = None
        if any(moving):
            off_dev = torch.as_tensor(offsets, device=dev)
            dt_dev = torch.as_tensor([j.req.get("dt", 0.25) if mv else 0.0 for j, mv in zip(jobs, moving)],
                                     dtype=torch.float64, device=dev)
            F_prev = torch.cat([j.forces if j.forces is not None else zeros3(j.n) for j in jobs]).contiguous()
            eng.verlet(off_dev, T, dt_dev, X, V, F_prev, 0)
        out = dict(forces=torch.empty((n, 3), dtype=torch.float64, device=dev),
                   charges=torch.empty(n, dtype=torch.float64, device=dev),
                   total_bond_orders=None, lone_pairs=None, bond_counts=None, qeq_state=QS)
        res = eng.evaluate_batch(X, T, offsets, cells, pbcs, charges, tol=self.tol, precision=self.precision,
                                 warm_start=self.warm_start, out=out)
        if any(md):
            if off_dev is None:
                off_dev = torch.as_tensor(offsets, device=dev)
                dt_dev = torch.zeros(ns, dtype=torch.float64, device=dev)
            kin = torch.zeros(ns, dtype=torch.float64, device=dev)
            eng.verlet(off_dev, T, dt_dev, X, V, out["forces"], 1, kin)
            kin = kin.cpu().numpy()
        fire_status = None
        if any(relax):
            if off_dev is None:
                off_dev = torch.as_tensor(offsets, device=dev)
            skip = torch.ones(1, dtype=torch.int32, device=dev)
            dummy = torch.zeros((1, 4), dtype=torch.float64, device=dev)
            FS = torch.cat([j.fire if r else dummy for j, r in zip(jobs, relax)]).contiguous()
            ST = torch.cat([j.status if r else skip for j, r in zip(jobs, relax)]).contiguous()
            eng.fire_step(off_dev, X, V, out["forces"], FS, ST, self.relax_tolerance, self.relax_iterations)
            fire_status = ST.cpu().numpy()
        self.atom_evals += n
        comps = res["components"]
        done = []
        for k, j in enumerate(jobs):
            lo, hi = offsets[k], offsets[k + 1]
            j.x = X[lo:hi]
            if V is not None and (md[k] or relax[k]):
                j.v = V[lo:hi]
            if relax[k]:
                j.fire = FS[k:k + 1]
                j.status = ST[k:k + 1]
            j.forces = out["forces"][lo:hi]
            j.qstate = QS[lo:hi]
            j.evals += 1
            j.potential.append(float(comps[k].sum()))
            if md[k]:
                j.kinetic.append(float(kin[k]))
            if relax[k]:
                finished_now = fire_status[k] != 0
            else:
                finished_now = j.evals > j.req.get("md_steps", 0)
            if res["errors"][k] is not None or finished_now:
                done.append((k, j))
        if not done:
            return
        F = out["forces"].cpu().numpy()
        Q = out["charges"].cpu().numpy()
        Xh = X.cpu().numpy() if any(md[k] or relax[k] for k, _ in done) else None
        Vh = V.cpu().numpy() if V is not None and any(md[k] for k, _ in done) else None
        now = self.clock()
        finished = set()
        for k, j in done:
            lo, hi = offsets[k], offsets[k + 1]
            r = Response(j.req["rid"], j.n, j.evals, float(comps[k].sum()), comps[k].copy(), F[lo:hi].copy(),
                         Q[lo:hi].copy(), j.arrival, now, res["errors"][k])
            if md[k]:
                r.positions = Xh[lo:hi].copy()
                r.velocities = Vh[lo:hi].copy()
                r.potential = np.array(j.potential)
                r.kinetic = np.array(j.kinetic)
            if relax[k]:
                # FIRE leaves a converged (or exhausted) system where it was evaluated
                r.positions = Xh[lo:hi].copy()
                r.potential = np.array(j.potential)
                r.converged = bool(fire_status[k] == 1)
            self.responses.append(r)
            self.used_bytes -= j.bytes
            finished.add(id(j))
        jobs[:] = [j for j in jobs if id(j) not in finished]

    # ------------------------------------------------------------------- run
    def idle(self):
        return not self.queue and not any(self.active.values())

    def run(self, requests, arrivals=None):
        """Serve ``requests`` arriving at ``arrivals`` (seconds from now; all at once if None).

        Returns the responses in completion order.
        """
        t0 = self.clock()
        order = np.argsort(arrivals, kind="stable") if arrivals is not None else np.arange(len(requests))
        times = np.zeros(len(requests)) if arrivals is None else np.asarray(arrivals, dtype=float)
        i = 0
        while i < len(order) or not self.idle():
            now = self.clock() - t0
            while i < len(order) and times[order[i]] <= now:
                self.submit(requests[order[i]], arrival=t0 + times[order[i]])
                i += 1
            if self.idle():
                wait = times[order[i]] - (self.clock() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
                continue
            self.step()
        return self.responses
