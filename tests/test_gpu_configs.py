"""The two large checking configurations on an MI355X.

* 3150-atom periodic CHO mixture (CHO.2008): the hand-written adjoint forces
  against the oracle's autograd forces on every atom, and against central
  differences (h = 1e-4 A) of the energy at fixed charges on 64 seeded
  coordinates.  Richardson extrapolation of the h and h/2 differences
  removes the O(h^2) truncation, so the comparison is tight.
* 24,360-atom anatase/water slab (TiOH.Kim2013): mixed-precision QEq (FP32
  sweeps, FP64 refinement) against the FP64 CG at tolerance 1e-10.  The FP64
  path is pinned to the oracle on the smaller TiOH cases; a dense oracle solve
  at this size is out of reach of a test.
"""

import numpy as np
import pytest

from oracle import reaxff as orc
from xreac_amd import ForceField, workloads

pytestmark = pytest.mark.gpu

_ENG = {}


def _engine(name):
    from xreac_amd.engine import Engine

    if name not in _ENG:
        _ENG[name] = Engine(ForceField.bundled(name))
    return _ENG[name]


def _dev(eng, x, s):
    import torch

    return (torch.as_tensor(np.asarray(x, dtype=np.float64), device=eng.device).contiguous(),
            torch.as_tensor(eng.type_ids(s), device=eng.device))


def test_cho_mixture_forces_vs_oracle():
    name = "ffield.reax.CHO.2008"
    s, x, cell = workloads.cho_mixture()
    assert len(s) == 3150
    eng = _engine(name)
    xd, td = _dev(eng, x, s)
    print("[c2] GPU evaluation", flush=True)
    gpu = eng.evaluate(xd, td, cell=cell, pbc=(True,) * 3, tol=1e-12)
    print("[c2] oracle evaluation", flush=True)
    ref = orc.evaluate(ForceField.bundled(name), s, x, cell=cell, pbc=True)
    e = float(gpu["components"].sum())
    assert abs(e - ref["energy"]) <= 1e-9 * abs(ref["energy"])
    np.testing.assert_allclose(gpu["forces"].cpu().numpy(), ref["forces"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(gpu["charges"].cpu().numpy(), ref["charges"], rtol=0, atol=1e-7)


def test_cho_mixture_forces_vs_central_differences():
    name = "ffield.reax.CHO.2008"
    s, x, cell = workloads.cho_mixture()
    x = np.asarray(x, dtype=np.float64)
    eng = _engine(name)
    xd, td = _dev(eng, x, s)
    base = eng.evaluate(xd, td, cell=cell, pbc=(True,) * 3, tol=1e-12)
    F = base["forces"].cpu().numpy().copy()
    q = base["charges"].cpu().numpy().copy()
    rng = np.random.default_rng(64)
    picks = [(int(a), int(k)) for a, k in zip(rng.choice(len(s), 64, replace=False), rng.integers(0, 3, 64))]

    def energy(y):
        yd, _ = _dev(eng, y, s)
        r = eng.evaluate(yd, td, cell=cell, pbc=(True,) * 3, charges=q, properties=False)
        return float(r["components"].sum())

    def central(a, k, h):
        xp, xm = x.copy(), x.copy()
        xp[a, k] += h
        xm[a, k] -= h
        return -(energy(xp) - energy(xm)) / (2 * h)

    worst = 0.0
    for a, k in picks:
        f1, f2 = central(a, k, 1e-4), central(a, k, 5e-5)
        rich = (4 * f2 - f1) / 3
        err = abs(rich - F[a, k])
        worst = max(worst, err / (1e-5 + 1e-6 * abs(F[a, k])))
        assert err <= 1e-5 + 1e-6 * abs(F[a, k]), (a, k, F[a, k], f1, f2, rich)
    print(f"[c2] worst FD error / tolerance = {worst:.3f}", flush=True)


def test_anatase_water_slab_mixed_vs_fp64():
    name = "ffield.reax.TiOH.Kim2013"
    s, x, cell = workloads.anatase_water_slab()
    assert len(s) == 24360
    eng = _engine(name)
    xd, td = _dev(eng, x, s)
    print("[c4] FP64 CG", flush=True)
    a = eng.evaluate(xd, td, cell=cell, pbc=(True,) * 3, tol=1e-10, precision="fp64")
    ea, fa, qa = float(a["components"].sum()), a["forces"].cpu().numpy().copy(), a["charges"].cpu().numpy().copy()
    print(f"[c4] mixed (FP64 took {a['qeq_iterations']} iterations)", flush=True)
    b = eng.evaluate(xd, td, cell=cell, pbc=(True,) * 3, tol=1e-10, precision="mixed")
    eb, fb, qb = float(b["components"].sum()), b["forces"].cpu().numpy(), b["charges"].cpu().numpy()
    dq, de, df = np.max(np.abs(qa - qb)), abs(ea - eb), np.max(np.abs(fa - fb))
    print(f"[c4] mixed vs FP64: max|dq| {dq:.2e} e, |dE| {de:.2e} kcal/mol ({de / abs(ea):.2e} rel), "
          f"max|dF| {df:.2e} kcal/mol/A; iterations {b['qeq_iterations']}", flush=True)
    assert dq <= 1e-7
    assert de <= 1e-9 * abs(ea)
    assert df <= 1e-6
