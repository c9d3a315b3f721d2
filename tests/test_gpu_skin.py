"""Skin reuse of device-built neighbor lists (xreac_options.neighbor_skin).

The reference's ASE calculator keeps a list with a per-atom skin and rebuilds
it only after an atom has moved more than the skin (src/xreac/ase.py:107-131,
neighbor_list_builds).  Here the kept list lives on the device and each call
filters it to the cutoff, so every result must be bitwise the result of a
fresh build; rebuilds must happen exactly when an atom moved more than the
skin, a type changed or the call's shape changed.
"""

import numpy as np
import pytest

from xreac_amd import ForceField, workloads

pytestmark = pytest.mark.gpu

HO = "ffield.reax.HO.2015"
SKIN = 0.3


def _engines():
    from xreac_amd.engine import Engine

    ff = ForceField.bundled(HO)
    return Engine(ff), Engine(ff)


def _same(a, b, what):
    np.testing.assert_array_equal(a["components"], b["components"], err_msg=what)
    np.testing.assert_array_equal(a["forces"].cpu().numpy(), b["forces"].cpu().numpy(), err_msg=what)
    np.testing.assert_array_equal(a["charges"].cpu().numpy(), b["charges"].cpu().numpy(), err_msg=what)
    assert a["num_edges"] == b["num_edges"] and a["num_bonds"] == b["num_bonds"], what


def _nve_case(name):
    if name == "cell":
        s, x, cell = workloads.water_cell(200, seed=2)  # 600 atoms, periodic
        return s, x, cell, (True, True, True)
    s, x = workloads.water_cluster()  # 192 atoms, nonperiodic
    return s, x, None, (False, False, False)


@pytest.mark.parametrize("name", ["cell", "cluster"])
def test_skin_reuse_along_nve_is_bitwise_fresh(name):
    import torch

    keep, fresh = _engines()
    s, x0, cell, pbc = _nve_case(name)
    n = len(s)
    dev = keep.device
    t = torch.as_tensor(keep.type_ids(s), device=dev).contiguous()
    masses = np.array([15.999 if e == "O" else 1.008 for e in s])
    X = torch.as_tensor(np.asarray(x0, dtype=np.float64), device=dev).contiguous()
    V = torch.as_tensor(workloads.maxwell_boltzmann(masses, 300.0, seed=3), device=dev).contiguous()
    off = torch.as_tensor(np.array([0, n], dtype=np.int32), device=dev)
    dt = torch.full((1,), 0.25, dtype=torch.float64, device=dev)
    F = None
    steps = 40
    for step in range(steps):
        if step:
            keep.verlet(off, t, dt, X, V, F, 0)
        a = keep.evaluate(X, t, cell=cell, pbc=pbc, neighbor_skin=0.5)
        b = fresh.evaluate(X, t, cell=cell, pbc=pbc)
        _same(a, b, f"{name} step {step}")
        F = a["forces"]
        if step:
            keep.verlet(off, t, dt, X, V, F, 1)
    builds = keep.neighbor_list_builds
    print(f"[skin] {name}: {n} atoms, {steps} steps of 0.25 fs from 300 K, skin 0.5 A, {builds} list builds",
          flush=True)
    # the lattice-built liquid cell heats quickly as its contacts relax, so it rebuilds more often
    assert 1 <= builds <= steps // 3
    assert fresh.neighbor_list_builds == 0


def test_skin_reuse_in_a_batch_is_bitwise_fresh():
    import torch

    keep, fresh = _engines()
    systems = [_nve_case("cell"), _nve_case("cluster")]
    xs = [np.asarray(x, dtype=np.float64) for _, x, _, _ in systems]
    ts = [keep.type_ids(s) for s, _, _, _ in systems]
    offsets = np.array([0, len(xs[0]), len(xs[0]) + len(xs[1])], dtype=np.int32)
    cells = np.stack([np.asarray(systems[0][2]), np.eye(3)])
    pbcs = np.array([[1, 1, 1], [0, 0, 0]], dtype=np.int32)
    dev = keep.device
    t = torch.as_tensor(np.concatenate(ts), device=dev).contiguous()
    rng = np.random.default_rng(5)
    base = np.concatenate(xs)
    for step in range(12):
        # small random moves (at most 0.05 A per call, so well inside the skin for several calls)
        pos = base + rng.uniform(-0.05, 0.05, size=base.shape) * (step > 0)
        base = pos
        X = torch.as_tensor(pos, device=dev).contiguous()
        a = keep.evaluate_batch(X, t, offsets, cells, pbcs, neighbor_skin=SKIN)
        b = fresh.evaluate_batch(X, t, offsets, cells, pbcs)
        assert all(e is None for e in a["errors"]), a["errors"]
        np.testing.assert_array_equal(a["components"], b["components"])
        np.testing.assert_array_equal(a["forces"].cpu().numpy(), b["forces"].cpu().numpy())
        np.testing.assert_array_equal(a["charges"].cpu().numpy(), b["charges"].cpu().numpy())
        assert a["num_edges"] == b["num_edges"]
    print(f"[skin] batch of 2: {keep.neighbor_list_builds} system list builds over 12 calls", flush=True)
    assert 2 <= keep.neighbor_list_builds < 12


def test_skin_rebuilds_exactly_when_needed():
    import torch

    keep, fresh = _engines()
    s, x0, cell, pbc = _nve_case("cell")
    dev = keep.device
    t0 = keep.type_ids(s)
    x = np.asarray(x0, dtype=np.float64).copy()

    def call(pos, types=t0, **kw):
        X = torch.as_tensor(pos, device=dev).contiguous()
        T = torch.as_tensor(types, device=dev).contiguous()
        a = keep.evaluate(X, T, cell=kw.get("cell", cell), pbc=pbc, neighbor_skin=kw.get("skin", SKIN))
        b = fresh.evaluate(X, T, cell=kw.get("cell", cell), pbc=pbc)
        _same(a, b, kw.get("what", ""))
        return keep.neighbor_list_builds

    assert call(x, what="first") == 1
    assert call(x, what="same positions") == 1
    y = x.copy()
    y[7] += [0.9 * SKIN, 0.0, 0.0]
    assert call(y, what="0.9 skin") == 1  # within the allowance: filtered, not rebuilt
    z = x.copy()
    z[7] += [0.0, 1.05 * SKIN, 0.0]
    assert call(z, what="1.05 skin") == 2  # beyond it (measured from the build positions)
    assert call(z, what="again") == 2
    w = z.copy()
    w[11] += np.asarray(cell)[0]  # the same configuration with one atom wrapped a cell over
    assert call(w, what="wrapped") == 3
    t1 = np.asarray(t0).copy()
    t1[[0, 1]] = t1[[1, 0]]  # two atoms swap species
    assert call(w, t1, what="types") == 4
    assert call(w, t1, skin=0.25, what="other skin") == 5
    assert call(w, t1, cell=np.asarray(cell) * 1.0001, what="other cell") == 6


def test_skin_option_errors():
    import torch

    from xreac_amd.engine import DeviceError

    keep, _ = _engines()
    s, x, cell, pbc = _nve_case("cluster")
    X = torch.as_tensor(np.asarray(x, dtype=np.float64), device=keep.device).contiguous()
    T = torch.as_tensor(keep.type_ids(s), device=keep.device).contiguous()
    for bad in (-0.1, float("nan"), float("inf")):
        with pytest.raises((ValueError, DeviceError), match="neighbor_skin must be a finite nonnegative number"):
            keep.evaluate(X, T, neighbor_skin=bad)
    with pytest.raises((ValueError, DeviceError), match="keeps no skin lists"):
        keep.evaluate(X, T, neighbor_skin=SKIN, no_alloc=True)
