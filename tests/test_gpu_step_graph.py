"""Recorded steps (a step of a fixed batch captured as a HIP graph and replayed).

No reference counterpart: the reference evaluates one system per host-driven
call.  Every replay must give bitwise the results of the same step run
unrecorded (verlet stage 0, QEq extrapolation, the batch evaluation, verlet
stage 1), step after step, with the unrecorded engine's own launches and
synchronisations between the replays.  A replay whose buffers overflow, or
whose QEq needs more launches than were recorded, must restore its inputs and
fall back to the unrecorded step with the same results.
"""

import numpy as np
import pytest

from xreac_amd import ForceField, workloads

pytestmark = pytest.mark.gpu

HO = "ffield.reax.HO.2015"


def _engines():
    import torch

    from xreac_amd.engine import Engine

    ff = ForceField.bundled(HO)
    return Engine(ff), Engine(ff, stream=torch.cuda.Stream(0))


def _mix():
    """A 60-atom cluster, a periodic 96-atom water cell and a charged 31-atom cluster."""
    s1, x1 = workloads.water_cluster_of(20, 11)
    s2, x2, c2 = workloads.water_cell(32, seed=12)
    s3, x3, q3 = workloads.ion_cluster(10, 1, 13)
    sym = list(s1) + list(s2) + list(s3)
    x = np.vstack([x1, x2, x3])
    offsets = np.array([0, len(s1), len(s1) + len(s2), len(sym)], dtype=np.int32)
    cells = np.stack([np.eye(3), np.asarray(c2), np.eye(3)])
    pbc = np.array([[0, 0, 0], [1, 1, 1], [0, 0, 0]], dtype=np.int32)
    charges = np.array([0.0, 0.0, q3])
    return sym, x, offsets, cells, pbc, charges


class _State:
    def __init__(self, eng, sym, x, v):
        import torch

        dev = eng.device
        f64 = dict(dtype=torch.float64, device=dev)
        n = len(sym)
        self.x = torch.as_tensor(x, **f64).contiguous()
        self.v = torch.as_tensor(v, **f64).contiguous()
        self.t = torch.as_tensor(eng.type_ids(sym), device=dev)
        self.out = eng.batch_outputs(n, qeq_state=True)
        self.out["forces"].zero_()
        self.qh = torch.zeros((3, n, 2), **f64)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 5e-8)])
def test_recorded_nve_matches_unrecorded(precision, tol):
    import torch

    eager, rec = _engines()
    ff = ForceField.bundled(HO)
    sym, x, offsets, cells, pbc, charges = _mix()
    ns = len(offsets) - 1
    masses = np.array([ff.atoms[a]["mass"] for a in sym])
    v = workloads.maxwell_boltzmann(masses, 300.0, seed=7)
    A, B = _State(eager, sym, x, v), _State(rec, sym, x, v)
    torch.cuda.synchronize()  # the states were made on the default stream; the recorded engine has its own
    dev = eager.device
    off_dev = torch.as_tensor(offsets, device=dev)
    dt_host = np.full(ns, 0.25)
    kin_a = torch.zeros(ns, dtype=torch.float64, device=dev)
    kin_b = torch.zeros(ns, dtype=torch.float64, device=dev)
    dt_b = torch.zeros(ns, dtype=torch.float64, device=dev)
    order_b = torch.zeros(ns, dtype=torch.int32, device=dev)
    step = rec.record_step(B.x, B.t, offsets, cells, pbc, charges, out=B.out, tol=tol, precision=precision,
                           warm_start=True, dt_fs=dt_b, velocities=B.v, kinetic=kin_b, qeq_order=order_b,
                           qeq_history=B.qh)
    steps = 20
    for e in range(steps):
        moving = e > 0
        dt = torch.as_tensor(dt_host if moving else np.zeros(ns), dtype=torch.float64, device=dev)
        order = torch.as_tensor(np.full(ns, min(e, 4) - 1 if moving else 0, dtype=np.int32), device=dev)
        # the unrecorded step, as the server runs it
        if moving:
            eager.verlet(off_dev, A.t, dt, A.x, A.v, A.out["forces"], 0)
            eager.qeq_extrapolate(off_dev, order, A.out["qeq_state"], A.qh)
        ra = eager.evaluate_batch(A.x, A.t, offsets, cells, pbc, charges, tol=tol, precision=precision,
                                  warm_start=True, out=A.out)
        eager.verlet(off_dev, A.t, dt, A.x, A.v, A.out["forces"], 1, kin_a)
        torch.cuda.synchronize()
        # the recorded one, its per-step inputs updated in place
        with torch.cuda.stream(rec.stream):
            dt_b.copy_(dt)
            order_b.copy_(order)
        rb = step.run()
        print(f"[step graph] {precision} step {e}: stats {step.stats()}", flush=True)
        assert all(err is None for err in ra["errors"]), ra["errors"]
        assert rb["errors"] == ra["errors"]
        np.testing.assert_array_equal(rb["components"], ra["components"])
        np.testing.assert_array_equal(rb["qeq_iterations"], ra["qeq_iterations"])
        for name in ("forces", "charges", "qeq_state"):
            assert torch.equal(B.out[name], A.out[name]), (e, name)
        assert torch.equal(B.x, A.x) and torch.equal(B.v, A.v), e
        assert torch.equal(kin_b, kin_a), e
    replays, fallbacks, recordings = step.stats()
    # recorded at the first run; recorded again with fewer QEq launches once
    # eight replays settled well inside the first recording's count
    assert (replays, fallbacks) == (steps - 1, 0) and recordings >= 1
    step.close()


def test_recorded_step_falls_back_on_overflow():
    import torch

    eager, rec = _engines()
    s, x = workloads.water_cluster_of(64, 21)
    A, B = _State(eager, s, x, np.zeros_like(x)), _State(rec, s, x, np.zeros_like(x))
    torch.cuda.synchronize()
    offsets = np.array([0, len(s)], dtype=np.int32)
    step = rec.record_step(B.x, B.t, offsets, out=B.out, tol=1e-10)
    step.run()
    assert step.stats() == (0, 0, 1)
    # compressed by 0.8 about the centroid: about twice the pairs within the
    # cutoff, past the recorded edge buffer
    c = x.mean(axis=0)
    xc = c + 0.8 * (x - c)
    for rep in range(2):
        with torch.cuda.stream(rec.stream):
            B.x.copy_(torch.as_tensor(xc, device=rec.device))
        A.x.copy_(torch.as_tensor(xc, device=eager.device))
        rb = step.run()
        ra = eager.evaluate_batch(A.x, A.t, offsets, tol=1e-10, out=A.out)
        torch.cuda.synchronize()
        print(f"[step graph] overflow rep {rep}: edges {rb['num_edges']} stats {step.stats()}", flush=True)
        assert ra["errors"] == [None] and rb["errors"] == [None]
        np.testing.assert_array_equal(rb["components"], ra["components"])
        assert torch.equal(B.out["forces"], A.out["forces"])
        assert torch.equal(B.out["charges"], A.out["charges"])
    # the first compressed run fell back and recorded again; the second replayed
    assert step.stats() == (1, 1, 2)
    step.close()


def test_recorded_step_falls_back_when_qeq_needs_more_launches():
    import torch

    eager, rec = _engines()
    s, x = workloads.water_cluster_of(40, 31)
    A, B = _State(eager, s, x, np.zeros_like(x)), _State(rec, s, x, np.zeros_like(x))
    torch.cuda.synchronize()
    offsets = np.array([0, len(s)], dtype=np.int32)
    # converge once, then eight warm solves: the recording sizes its QEq
    # launches from those (a few launches each)
    for k in range(9):
        rec.evaluate_batch(B.x, B.t, offsets, tol=1e-10, warm_start=k > 0, out=B.out)
    step = rec.record_step(B.x, B.t, offsets, out=B.out, tol=1e-10, warm_start=True)
    step.run()
    step.run()
    assert step.stats() == (1, 0, 1)
    # a cold guess needs far more launches than were recorded
    with torch.cuda.stream(rec.stream):
        B.out["qeq_state"].zero_()
    rb = step.run()
    ra = eager.evaluate_batch(A.x, A.t, offsets, tol=1e-10, warm_start=True, out=A.out)  # zero guess too
    torch.cuda.synchronize()
    print(f"[step graph] cold guess: iterations {rb['qeq_iterations']} stats {step.stats()}", flush=True)
    assert step.stats() == (1, 1, 2)
    assert ra["errors"] == [None] and rb["errors"] == [None]
    np.testing.assert_array_equal(rb["components"], ra["components"])
    np.testing.assert_array_equal(rb["qeq_iterations"], ra["qeq_iterations"])
    for name in ("forces", "charges", "qeq_state"):
        assert torch.equal(B.out[name], A.out[name]), name
    step.close()
