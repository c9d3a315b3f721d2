"""Device integrators against their CPU restatements driven by the FP64 oracle.

Velocity Verlet (no reference counterpart) and FIRE (calculator.py:213-246)
run on the GPU over a batch; each system must follow the oracle trajectory.
"""

import numpy as np
import pytest

from oracle import dynamics as od
from oracle import reaxff as orc
from tests import cases
from xreac_amd import ForceField, workloads

pytestmark = pytest.mark.gpu

HO = "ffield.reax.HO.2015"
_ENG = {}


def _engine():
    from xreac_amd.engine import Engine

    if "e" not in _ENG:
        _ENG["e"] = Engine(ForceField.bundled(HO))
    return _ENG["e"]


def _cluster(molecules, seed):
    s, x = workloads.water_cluster(molecules=molecules, radius=4.0, seed=seed)
    return s, x


def test_verlet_batch_matches_oracle():
    import torch

    eng = _engine()
    ff = ForceField.bundled(HO)
    s1, x1 = _cluster(4, 3)
    s2, x2 = cases.water_cases()["distorted_dimer"]
    masses = np.array([ff.atoms[a]["mass"] for a in s1])
    v1 = workloads.maxwell_boltzmann(masses, 300.0, seed=5)
    dev = eng.device
    x = torch.as_tensor(np.vstack([x1, x2]), device=dev).contiguous()
    v = torch.as_tensor(np.vstack([v1, np.zeros((len(s2), 3))]), device=dev).contiguous()
    t = torch.as_tensor(np.concatenate([eng.type_ids(s1), eng.type_ids(s2)]), device=dev)
    off_h = np.array([0, len(s1), len(s1) + len(s2)], dtype=np.int32)
    off = torch.as_tensor(off_h, device=dev)
    dt = torch.as_tensor([0.25, 0.0], dtype=torch.float64, device=dev)
    kin = torch.zeros(2, dtype=torch.float64, device=dev)
    out = eng.batch_outputs(len(t), qeq_state=True)
    steps = 20
    res = eng.evaluate_batch(x, t, off_h, tol=1e-12, out=out)
    pot = [res["components"][0].sum()]
    kins = []
    for k in range(steps):
        eng.verlet(off, t, dt, x, v, out["forces"], 0)
        res = eng.evaluate_batch(x, t, off_h, tol=1e-12, out=out, warm_start=True)
        assert all(e is None for e in res["errors"]), res["errors"]
        eng.verlet(off, t, dt, x, v, out["forces"], 1, kin)
        pot.append(res["components"][0].sum())
        kins.append(kin.cpu().numpy()[0])
    print(f"[verlet] {steps} steps done", flush=True)

    def force_fn(xx):
        r = orc.evaluate(ff, s1, xx, qeq_tolerance=1e-12) if "qeq_tolerance" in orc.evaluate.__code__.co_varnames \
            else orc.evaluate(ff, s1, xx)
        return r["energy"], r["forces"]

    xr, vr, pr, kr = od.velocity_verlet(force_fn, x1, v1, masses, 0.25, steps)
    xg = x.cpu().numpy()
    vg = v.cpu().numpy()
    np.testing.assert_allclose(xg[: len(s1)], xr, rtol=0, atol=1e-8)
    np.testing.assert_allclose(vg[: len(s1)], vr, rtol=0, atol=1e-10)
    np.testing.assert_allclose(np.array(pot), pr, rtol=0, atol=1e-6)
    np.testing.assert_allclose(np.array(kins), kr[1:], rtol=0, atol=1e-8)
    # the system with dt = 0 has not moved
    np.testing.assert_array_equal(xg[len(s1):], np.asarray(x2))
    # short NVE segment: total energy stays within a small band
    total = np.array(pot[1:]) + np.array(kins)
    assert np.ptp(total) < 0.05, np.ptp(total)


def test_fire_batch_matches_oracle():
    import torch

    eng = _engine()
    ff = ForceField.bundled(HO)
    systems = [cases.water_cases()["distorted_dimer"], _cluster(3, 8)]
    dev = eng.device
    x = torch.as_tensor(np.vstack([sx for _, sx in systems]), device=dev).contiguous()
    t = torch.as_tensor(np.concatenate([eng.type_ids(ss) for ss, _ in systems]), device=dev)
    off_h = np.cumsum([0] + [len(ss) for ss, _ in systems]).astype(np.int32)
    off = torch.as_tensor(off_h, device=dev)
    v = torch.zeros_like(x)
    state = torch.tensor([[0.02, 0.1, 0.0, 0.0]] * len(systems), dtype=torch.float64, device=dev)
    status = torch.zeros(len(systems), dtype=torch.int32, device=dev)
    out = eng.batch_outputs(len(t), qeq_state=True)
    iters = 12
    traj = [x.cpu().numpy().copy()]
    for k in range(iters):
        eng.evaluate_batch(x, t, off_h, tol=1e-12, out=out, warm_start=k > 0)
        eng.fire_step(off, x, v, out["forces"], state, status, 1e-4, 1000)
        traj.append(x.cpu().numpy().copy())
    print(f"[fire] {iters} iterations done", flush=True)
    for q, (ss, sx) in enumerate(systems):
        _, _, _, ref = od.fire(lambda xx: orc.evaluate(ff, ss, xx)["forces"], sx, 1e-4, iters)
        for k in range(iters + 1):
            np.testing.assert_allclose(traj[k][off_h[q]:off_h[q + 1]], ref[k], rtol=0, atol=1e-9,
                                       err_msg=f"system {q} iteration {k}")
    assert (status.cpu().numpy() == 0).all()


def test_calculator_relax_native_matches_oracle_fire():
    """Calculator.relax(backend="native") (calculator.py:167-246) on the reference's ZnO dimer."""
    from xreac_amd import Calculator

    ff = ForceField.bundled("ffield.reax.ZnOH.2010")
    calc = Calculator(ff, qeq_tolerance=1e-12)
    s, x = cases.ZNO["zno"]
    rel = calc.relax(s, x, backend="native")
    assert rel.converged and rel.message == "Fixed-charge force tolerance reached"
    xr, converged, it, _ = od.fire(lambda y: orc.evaluate(ff, s, y)["forces"], x, 1e-4, 500)
    assert converged and rel.iterations == it
    np.testing.assert_allclose(rel.positions, xr, rtol=0, atol=1e-7)
    assert np.max(np.abs(rel.evaluation.forces)) <= 1e-4


def test_qeq_extrapolation_is_exact_on_polynomials():
    """The degree-k guess reproduces any polynomial of degree <= k sampled at equal steps (integer
    coefficients: exact in binary), leaves order 0 untouched and shifts the history by one."""
    import torch

    eng = _engine()
    dev = eng.device
    sizes = [5, 7, 3, 9, 4]
    order = np.array([0, 1, 2, 3, 3], dtype=np.int32)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n = int(off[-1])
    rng = np.random.default_rng(3)
    # per atom and right-hand side: p(t) = c0 + c1 t + c2 t^2 + c3 t^3 with small integer c, the
    # coefficients above each system's order zero; solutions at t = 0 (current), -1, -2, -3 (history)
    coef = rng.integers(-9, 10, size=(n, 2, 4)).astype(np.float64)
    deg = np.repeat(order, sizes)
    for d in range(4):
        coef[deg < d, :, d] = 0.0

    def p(t):
        return coef[..., 0] + coef[..., 1] * t + coef[..., 2] * t**2 + coef[..., 3] * t**3

    state = torch.as_tensor(p(0.0), device=dev).contiguous()
    hist = torch.as_tensor(np.stack([p(-1.0), p(-2.0), p(-3.0)]), device=dev).contiguous()
    eng.qeq_extrapolate(torch.as_tensor(off, device=dev), torch.as_tensor(order, device=dev), state, hist)
    got = state.cpu().numpy()
    np.testing.assert_array_equal(got, p(1.0))
    h = hist.cpu().numpy()
    np.testing.assert_array_equal(h[0], p(0.0))
    np.testing.assert_array_equal(h[1], p(-1.0))
    np.testing.assert_array_equal(h[2], p(-2.0))
    # a degree-k guess is not exact one degree up: the cubic term shows up with coefficient 6 = 3!
    coef[:, :, 3] = 1.0
    state = torch.as_tensor(p(0.0), device=dev).contiguous()
    hist = torch.as_tensor(np.stack([p(-1.0), p(-2.0), p(-3.0)]), device=dev).contiguous()
    two = np.full(len(sizes), 2, dtype=np.int32)
    eng.qeq_extrapolate(torch.as_tensor(off, device=dev), torch.as_tensor(two, device=dev), state, hist)
    np.testing.assert_array_equal(state.cpu().numpy(), p(1.0) - 6.0)
