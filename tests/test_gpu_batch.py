"""Batched evaluation on an MI355X: each system of a batch gives bitwise the
result it gives alone, failures stay with their system, and the
mixed-precision and warm-started QEq solves meet the FP64 parity bars."""

import numpy as np
import pytest

from oracle import reaxff as orc
from tests import cases
from xreac_amd import ForceField, workloads

pytestmark = pytest.mark.gpu

HO = "ffield.reax.HO.2015"
E_REL = 1e-9
F_ABS = 1e-6
Q_ABS = 1e-7


def _ho_systems():
    out = []
    for name, (s, x) in cases.water_cases().items():
        out.append((name, s, x, None, None, 0.0))
    for name, (s, x, cell, pbc) in cases.periodic_cases(include_bulk=True).items():
        out.append((name, s, x, cell, pbc, 0.0))
    for name, (s, x, cell, pbc, q) in cases.charged_cases().items():
        out.append((name, s, x, cell, pbc if cell is not None else None, q))
    for name, (ffn, s, x, cell, pbc) in cases.small_cell_cases().items():
        if ffn == HO:
            out.append((name, s, x, cell, pbc, 0.0))
    s, x = workloads.water_cluster()
    out.append(("cluster192", s, x, None, None, 0.0))
    s, x, cell = workloads.water_box(side=4, seed=5)
    out.append(("box192", s, x, cell, (True, True, True), 0.0))
    return out


SYSTEMS = _ho_systems()
_ENGINES = {}


def _engine():
    from xreac_amd.engine import Engine

    if HO not in _ENGINES:
        _ENGINES[HO] = Engine(ForceField.bundled(HO))
    return _ENGINES[HO]


def _pack(eng, systems):
    import torch

    xs, ts, offsets, cells, pbcs, qs = [], [], [0], [], [], []
    for _, s, x, cell, pbc, q in systems:
        xs.append(np.asarray(x, dtype=np.float64))
        ts.append(eng.type_ids(s))
        offsets.append(offsets[-1] + len(s))
        cells.append(np.eye(3) if cell is None else np.asarray(cell, dtype=np.float64).reshape(3, 3))
        pbcs.append([0, 0, 0] if cell is None else [int(bool(v)) for v in (pbc if pbc is not None else (1, 1, 1))])
        qs.append(q)
    x = torch.as_tensor(np.concatenate(xs), device=eng.device).contiguous()
    t = torch.as_tensor(np.concatenate(ts), device=eng.device).contiguous()
    return x, t, np.array(offsets), np.array(cells), np.array(pbcs), np.array(qs, dtype=np.float64)


def _alone(eng, system, precision="fp64"):
    import torch

    _, s, x, cell, pbc, q = system
    xd = torch.as_tensor(np.asarray(x, dtype=np.float64), device=eng.device).contiguous()
    td = torch.as_tensor(eng.type_ids(s), device=eng.device)
    return eng.evaluate(xd, td, cell=cell, pbc=pbc if pbc is not None else (True,) * 3 if cell is not None else
                        (False,) * 3, total_charge=q, tol=1e-12, precision=precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_is_bitwise_alone(precision):
    eng = _engine()
    print(f"[batch] {len(SYSTEMS)} systems, {precision}", flush=True)
    x, t, off, cells, pbcs, qs = _pack(eng, SYSTEMS)
    res = eng.evaluate_batch(x, t, off, cells, pbcs, qs, tol=1e-12, precision=precision, properties=True)
    assert all(e is None for e in res["errors"]), res["errors"]
    forces = res["forces"].cpu().numpy()
    charges = res["charges"].cpu().numpy()
    tbo = res["total_bond_orders"].cpu().numpy()
    for k, system in enumerate(SYSTEMS):
        print(f"[batch] alone: {system[0]}", flush=True)
        a = _alone(eng, system, precision)
        lo, hi = off[k], off[k + 1]
        np.testing.assert_array_equal(res["components"][k], a["components"], err_msg=system[0])
        np.testing.assert_array_equal(forces[lo:hi], a["forces"].cpu().numpy(), err_msg=system[0])
        np.testing.assert_array_equal(charges[lo:hi], a["charges"].cpu().numpy(), err_msg=system[0])
        np.testing.assert_array_equal(tbo[lo:hi], a["total_bond_orders"].cpu().numpy(), err_msg=system[0])
        assert res["qeq_iterations"][k] == a["qeq_iterations"], system[0]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_with_both_cg_block_shapes(precision):
    # Systems below 8,192 atoms take the small CG block shape and larger ones
    # the large shape, so this batch runs two iteration launches per CG
    # iteration, with the large system's blocks numbered after both small
    # systems' although it sits between them.
    eng = _engine()
    s, x, cell = workloads.water_box(side=14, seed=7)
    assert len(s) >= 8192
    systems = [SYSTEMS[0], ("box8232", s, x, cell, (True, True, True), 0.0), SYSTEMS[-1]]
    print(f"[batch] both shapes, {precision}", flush=True)
    x, t, off, cells, pbcs, qs = _pack(eng, systems)
    res = eng.evaluate_batch(x, t, off, cells, pbcs, qs, tol=1e-12, precision=precision)
    assert all(e is None for e in res["errors"]), res["errors"]
    forces = res["forces"].cpu().numpy()
    charges = res["charges"].cpu().numpy()
    for k, system in enumerate(systems):
        print(f"[batch] alone: {system[0]}", flush=True)
        a = _alone(eng, system, precision)
        lo, hi = off[k], off[k + 1]
        np.testing.assert_array_equal(res["components"][k], a["components"], err_msg=system[0])
        np.testing.assert_array_equal(forces[lo:hi], a["forces"].cpu().numpy(), err_msg=system[0])
        np.testing.assert_array_equal(charges[lo:hi], a["charges"].cpu().numpy(), err_msg=system[0])
        assert res["qeq_iterations"][k] == a["qeq_iterations"], system[0]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_beyond_16_bit_atom_indices(precision):
    # 67,854 atoms: global atom indices no longer fit 16 bits, but every
    # system's do, so the CG sweeps read 2-byte columns relative to each
    # system's first atom; systems placed past atom 65,536 must still give
    # bitwise their results alone.
    eng = _engine()
    systems = []
    for side, seed in ((21, 8), (16, 10), (21, 9)):
        s, x, cell = workloads.water_box(side=side, seed=seed)
        systems.append((f"box{len(s)}_{seed}", s, x, cell, (True, True, True), 0.0))
    systems.append(SYSTEMS[-1])
    x, t, off, cells, pbcs, qs = _pack(eng, systems)
    assert off[-1] > 65536 and off[2] < 65536 < off[3]
    print(f"[batch] {off[-1]} atoms, {precision}", flush=True)
    res = eng.evaluate_batch(x, t, off, cells, pbcs, qs, tol=1e-12, precision=precision)
    assert all(e is None for e in res["errors"]), res["errors"]
    forces = res["forces"].cpu().numpy()
    charges = res["charges"].cpu().numpy()
    for k, system in enumerate(systems):
        print(f"[batch] alone: {system[0]}", flush=True)
        a = _alone(eng, system, precision)
        lo, hi = off[k], off[k + 1]
        np.testing.assert_array_equal(res["components"][k], a["components"], err_msg=system[0])
        np.testing.assert_array_equal(forces[lo:hi], a["forces"].cpu().numpy(), err_msg=system[0])
        np.testing.assert_array_equal(charges[lo:hi], a["charges"].cpu().numpy(), err_msg=system[0])
        assert res["qeq_iterations"][k] == a["qeq_iterations"], system[0]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_of_many_small_systems_is_bitwise_alone(precision):
    # More than 1,024 small-system CG blocks: the batch runs them as 256-thread
    # blocks while each system alone runs 512-thread ones; the liquid cell of
    # 1,650 atoms has more than 64 blocks, so its partial sums take the
    # strided reduction.  Results must not depend on the shape.
    eng = _engine()
    clusters = [workloads.water_cluster_of(256, seed) for seed in range(6)]
    systems = [(f"cluster768_{k % 6}_{k}", *clusters[k % 6], None, None, 0.0) for k in range(36)]
    s, x, cell = workloads.water_cell(550, seed=1)
    systems.insert(17, ("cell1650", s, x, cell, (True, True, True), 0.0))
    systems.append(SYSTEMS[-1])
    x, t, off, cells, pbcs, qs = _pack(eng, systems)
    assert sum(-(-(hi - lo) // 24) for lo, hi in zip(off[:-1], off[1:])) > 1100
    print(f"[batch] {len(systems)} small systems, {off[-1]} atoms, {precision}", flush=True)
    res = eng.evaluate_batch(x, t, off, cells, pbcs, qs, tol=1e-12, precision=precision)
    assert all(e is None for e in res["errors"]), res["errors"]
    forces = res["forces"].cpu().numpy()
    charges = res["charges"].cpu().numpy()
    for k in (0, 17, 30, len(systems) - 1):
        system = systems[k]
        print(f"[batch] alone: {system[0]}", flush=True)
        a = _alone(eng, system, precision)
        lo, hi = off[k], off[k + 1]
        np.testing.assert_array_equal(res["components"][k], a["components"], err_msg=system[0])
        np.testing.assert_array_equal(forces[lo:hi], a["forces"].cpu().numpy(), err_msg=system[0])
        np.testing.assert_array_equal(charges[lo:hi], a["charges"].cpu().numpy(), err_msg=system[0])
        assert res["qeq_iterations"][k] == a["qeq_iterations"], system[0]


def test_batch_failure_stays_with_its_system():
    eng = _engine()
    bad = ("coincident", ["O", "H"], [[0.0, 0, 0], [12.0, 0, 0]], np.diag([12.0] * 3), (True, True, True), 0.0)
    systems = [SYSTEMS[0], bad, SYSTEMS[-1]]
    x, t, off, cells, pbcs, qs = _pack(eng, systems)
    res = eng.evaluate_batch(x, t, off, cells, pbcs, qs, tol=1e-12)
    assert res["errors"][0] is None and res["errors"][2] is None
    assert "Coincident" in res["errors"][1]
    forces = res["forces"].cpu().numpy()
    for k in (0, 2):
        a = _alone(eng, systems[k])
        np.testing.assert_array_equal(forces[off[k]:off[k + 1]], a["forces"].cpu().numpy())
        np.testing.assert_array_equal(res["components"][k], a["components"])


def test_batch_nonconvergence_reported_per_system():
    eng = _engine()
    # systems larger than one preconditioner block (a monomer's block is exact and converges at once)
    big = [sy for sy in SYSTEMS if sy[0] in ("cluster192", "box192", "bulk_water_192")]
    x, t, off, cells, pbcs, qs = _pack(eng, big)
    res = eng.evaluate_batch(x, t, off, cells, pbcs, qs, max_iter=1)
    assert all("did not converge" in e for e in res["errors"])


@pytest.mark.parametrize("name", ["cluster192", "box192", "charged_water_4A"])
def test_mixed_precision_parity(name):
    eng = _engine()
    system = [sy for sy in SYSTEMS if sy[0] == name][0]
    _, s, x, cell, pbc, q = system
    print(f"[mixed] {name}", flush=True)
    gpu = _alone(eng, system, "mixed")
    ref = orc.evaluate(ForceField.bundled(HO), s, x, cell=cell, pbc=pbc, total_charge=q)
    e = float(gpu["components"].sum())
    assert abs(e - ref["energy"]) <= E_REL * max(1.0, abs(ref["energy"]))
    np.testing.assert_allclose(gpu["forces"].cpu().numpy(), ref["forces"], rtol=0, atol=F_ABS)
    np.testing.assert_allclose(gpu["charges"].cpu().numpy(), ref["charges"], rtol=0, atol=Q_ABS)


def test_warm_start_reuses_the_solution():
    import torch

    eng = _engine()
    systems = [sy for sy in SYSTEMS if sy[0] in ("cluster192", "box192")]
    x, t, off, cells, pbcs, qs = _pack(eng, systems)
    out = eng.batch_outputs(int(off[-1]), qeq_state=True)
    cold = eng.evaluate_batch(x, t, off, cells, pbcs, qs, tol=1e-10, out=out)
    q_cold = cold["charges"].cpu().numpy().copy()
    f_cold = cold["forces"].cpu().numpy().copy()
    it_cold = cold["qeq_iterations"].copy()
    again = eng.evaluate_batch(x, t, off, cells, pbcs, qs, tol=1e-10, out=out, warm_start=True)
    assert (again["qeq_iterations"] <= 2).all(), again["qeq_iterations"]
    np.testing.assert_allclose(again["charges"].cpu().numpy(), q_cold, rtol=0, atol=1e-9)
    np.testing.assert_allclose(again["forces"].cpu().numpy(), f_cold, rtol=0, atol=1e-7)
    # a small displacement converges from the previous charges in fewer iterations
    rng = np.random.default_rng(11)
    moved = x + torch.as_tensor(rng.normal(0, 0.01, tuple(x.shape)), device=eng.device)
    warm = eng.evaluate_batch(moved.contiguous(), t, off, cells, pbcs, qs, tol=1e-10, out=out, warm_start=True)
    q_warm = warm["charges"].cpu().numpy().copy()
    it_warm = warm["qeq_iterations"].copy()
    fresh = eng.evaluate_batch(moved.contiguous(), t, off, cells, pbcs, qs, tol=1e-10)
    assert (it_warm < it_cold).all(), (it_warm, it_cold)
    np.testing.assert_allclose(q_warm, fresh["charges"].cpu().numpy(), rtol=0, atol=Q_ABS)
