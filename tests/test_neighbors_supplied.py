"""Caller-supplied neighbor lists (calculator.py:113-115, neighbors.py:101-131).

Host validation runs on the CPU with the reference's messages
(tests/test_neighbors.py:293-314); the device evaluation with a supplied list,
with and without a skin, is compared with the oracle on the GPU.
"""

import numpy as np
import pytest

from oracle import reaxff as orc
from tests import cases
from xreac_amd import Calculator, ForceField
from xreac_amd.geometry import canonical_neighbors, fit_device_range

ZNOH = "ffield.reax.ZnOH.2010"


@pytest.mark.parametrize(
    "neighbors,message",
    [
        (([0], [1]), "tuple"),
        (([0], [1, 0], [[0, 0, 0]]), "shape"),
        (([0], [1], [[0, 0]]), "shape"),
        (([0.0, 1.0], [1, 0], [[0, 0, 0], [0, 0, 0]]), "integers"),
        (([0, 1], [1, 0], [[0.0, 0, 0], [0, 0, 0]]), "integers"),
        (([-1, 1], [1, 0], [[0, 0, 0], [0, 0, 0]]), "out of range"),
        (([0, 1], [2, 0], [[0, 0, 0], [0, 0, 0]]), "out of range"),
        (([0, 1], [1, 0], [[1, 0, 0], [-1, 0, 0]]), "nonperiodic"),
        (([0], [0], [[0, 0, 0]]), "self neighbors"),
        (([0, 0, 1], [1, 1, 0], [[0, 0, 0]] * 3), "Duplicate"),
        (([0], [1], [[0, 0, 0]]), "both directions"),
    ],
)
def test_supplied_neighbor_validation(neighbors, message):
    # validation happens on the host, before any device work
    with pytest.raises(ValueError, match=message):
        Calculator(ForceField.bundled(ZNOH)).evaluate(["Zn", "O"], [[0, 0, 0], [1.9, 0, 0]], neighbors=neighbors)


@pytest.mark.parametrize("index_dtype", [np.int64, np.uint64, np.int32])
def test_supplied_order_and_reverse(index_dtype):
    """Canonical (i, j, Sx, Sy, Sz) order whatever the input order (neighbors.py:63-98)."""
    i = np.array([1, 0, 0, 1, 0, 1], dtype=index_dtype)
    j = 1 - i
    shifts = np.array([[3, 0, 0], [0, 0, 0], [-3, 0, 0], [0, 0, 0], [-4, 0, 0], [4, 0, 0]])
    row_ptr, jj, S = canonical_neighbors((i, j, shifts), 2, np.eye(3), True)
    ii = np.repeat(np.arange(2), np.diff(row_ptr))
    keys = list(zip(ii, jj, *S.T))
    assert keys == sorted(zip(i.astype(int), j.astype(int), *shifts.T))
    with pytest.raises(ValueError, match="both directions"):
        canonical_neighbors((i[:-1], j[:-1], shifts[:-1]), 2, np.eye(3), True)
    with pytest.raises(ValueError, match="Duplicate"):
        canonical_neighbors((np.r_[i, i[:1]], np.r_[j, j[:1]], np.concatenate((shifts, shifts[:1]))), 2, np.eye(3),
                            True)
    # shifts of any size are ordered like the reference's overflow-safe fallback (tests/test_neighbors.py:137-154)
    for big in (2**40, 2**62 - 8):
        wide = np.where(shifts != 0, np.sign(shifts) * (big + np.abs(shifts)), 0)
        row_ptr, jj, S = canonical_neighbors((i, j, wide), 2, np.eye(3), True)
        ii = np.repeat(np.arange(2), np.diff(row_ptr))
        assert list(zip(ii, jj, *S.T)) == sorted(zip(i.astype(int), j.astype(int), *wide.T))


def test_fit_device_range_keeps_edge_vectors():
    """Whole-cell moves of far atoms keep every edge vector; edges far beyond the cutoff are dropped."""
    ff = ForceField.bundled("ffield.reax.HO.2015")
    rc = ff.general[12]
    s, x = cases.water_cases()["distorted_dimer"]
    cell = np.diag([12.0, 13.0, 14.0])
    base = canonical_neighbors(orc.build_neighbors(x, rc, cell, True), len(x), cell, True)
    x_far = x.copy()
    x_far[4] += 1000 * cell[0] - 700 * cell[2]  # one atom unwrapped by many cells
    w = np.array([1000, 0, -700])
    i0 = np.repeat(np.arange(len(x)), np.diff(base[0]))
    S_far = base[2] - np.where(base[1][:, None] == 4, w, 0) + np.where(i0[:, None] == 4, w, 0)
    extra_i, extra_j = [0, 1, 2, 2], [1, 0, 2, 2]
    extra_S = [[2**40, 0, 0], [-(2**40), 0, 0], [0, 2**62 - 5, 0], [0, -(2**62 - 5), 0]]
    supplied = (np.r_[i0, extra_i], np.r_[base[1], extra_j], np.vstack((S_far, extra_S)))
    csr = canonical_neighbors(supplied, len(x), cell, True)
    y, fitted = fit_device_range(x_far, cell, True, rc, csr)
    np.testing.assert_array_equal(fitted[0], base[0])
    np.testing.assert_array_equal(fitted[1], base[1])
    assert fitted[2].dtype == np.int32 and np.max(np.abs(fitted[2])) <= 1
    rows = np.repeat(np.arange(len(x)), np.diff(base[0]))
    v0 = x[base[1]] - x[rows] + base[2] @ cell
    v1 = y[fitted[1]] - y[rows] + fitted[2] @ cell
    np.testing.assert_allclose(v1, v0, rtol=0, atol=1e-9)
    # in range: nothing moves
    y2, same = fit_device_range(x, cell, True, rc, base)
    assert y2 is x or np.array_equal(y2, x)
    np.testing.assert_array_equal(same[2], base[2])
    # a chain whose cell is so thin that >511 images lie within the cutoff cannot be packed
    k = np.arange(1, 1001)
    chain = (np.zeros(2000, dtype=int), np.zeros(2000, dtype=int),
             np.column_stack((np.r_[k, -k], np.zeros(2000, dtype=int), np.zeros(2000, dtype=int))))
    thin = np.diag([0.009, 12.0, 12.0])
    csr = canonical_neighbors(chain, 1, thin, [True, False, False])
    with pytest.raises(ValueError, match="511"):
        fit_device_range(np.zeros((1, 3)), thin, [True, False, False], rc, csr)


def _cases():
    out = {}
    for name, (ffn, s, x, cell, pbc) in cases.small_cell_cases().items():
        out["small_" + name] = (ffn, s, x, cell, pbc)
    out["periodic_carbon_chain"] = cases.periodic_carbon_chain()
    s, x, cell, pbc = cases.periodic_cases(include_bulk=True)["bulk_water_192"]
    out["bulk_water_192"] = ("ffield.reax.HO.2015", s, x, cell, pbc)
    s, x = cases.water_cases()["distorted_dimer"]
    out["water_dimer"] = ("ffield.reax.HO.2015", s, x, None, None)
    return out


CASES = _cases()


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("skin", [0.0, 1.0])
def test_supplied_list_on_gpu(name, skin):
    ffn, s, x, cell, pbc = CASES[name]
    ff = ForceField.bundled(ffn)
    print(f"[supplied] {name} skin={skin}", flush=True)
    supplied = orc.build_neighbors(np.asarray(x, dtype=np.float64), ff.general[12] + skin, cell, pbc)
    calc = Calculator(ff, qeq_tolerance=1e-12, max_expanded_atoms=10**9)
    gpu = calc.evaluate(s, x, cell=cell, pbc=pbc, neighbors=supplied)
    assert gpu.neighbor_backend == "provided" and gpu.cell_repetitions == (1, 1, 1)
    ref = orc.evaluate(ff, s, x, cell=cell, pbc=pbc, neighbors=supplied)
    assert abs(gpu.energy - ref["energy"]) <= 1e-9 * max(1.0, abs(ref["energy"]))
    np.testing.assert_allclose(gpu.forces, ref["forces"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(gpu.charges, ref["charges"], rtol=0, atol=1e-7)
    np.testing.assert_array_equal(gpu.bond_counts, ref["bond_counts"])
    built = calc.evaluate(s, x, cell=cell, pbc=pbc)
    assert abs(gpu.energy - built.energy) <= 1e-9 * max(1.0, abs(built.energy))
    np.testing.assert_allclose(gpu.forces, built.forces, rtol=0, atol=1e-8)


@pytest.mark.gpu
def test_supplied_zinc_chain_images():
    """Self-image list of a one-atom Zn chain (tests/test_neighbors.py:275-290)."""
    calc = Calculator(ForceField.bundled(ZNOH))
    shifts = np.array([[k, 0, 0] for k in (-4, -3, -2, -1, 1, 2, 3, 4)])
    neighbors = (np.zeros(8, dtype=int), np.zeros(8, dtype=int), shifts)
    actual = calc.evaluate(["Zn"], [[0, 0, 0]], cell=[2.5, 12, 12], pbc=[True, False, False], neighbors=neighbors)
    expected = calc.evaluate(["Zn"], [[0, 0, 0]], cell=[2.5, 12, 12], pbc=[True, False, False])
    assert abs(actual.energy - expected.energy) < 1e-10
    assert actual.bond_counts.tolist() == [2]


@pytest.mark.gpu
def test_reuse_supplied_skin_list_for_displacements():
    """One skin list reused over small displacements (tests/test_neighbors.py:316-332)."""
    ffn, s, x, cell, pbc = CASES["small_water_4A"]
    ff = ForceField.bundled(ffn)
    supplied = orc.build_neighbors(np.asarray(x, dtype=np.float64), ff.general[12] + 1.0, cell, pbc)
    originals = tuple(np.array(a, copy=True) for a in supplied)
    calc = Calculator(ff, qeq_tolerance=1e-12)
    for shift in (0.0, 0.03, -0.02):
        y = np.array(x, dtype=np.float64)
        y[1, 0] += shift
        actual = calc.evaluate(s, y, cell=cell, pbc=pbc, neighbors=supplied)
        expected = calc.evaluate(s, y, cell=cell, pbc=pbc)
        assert abs(actual.energy - expected.energy) <= 1e-9 * abs(expected.energy)
        np.testing.assert_allclose(actual.forces, expected.forces, rtol=0, atol=1e-8)
    for original, array in zip(originals, supplied):
        np.testing.assert_array_equal(original, array)


@pytest.mark.gpu
def test_huge_shifts_and_far_atoms_on_gpu():
    """Supplied shifts up to 2**62 and atoms unwrapped by 1000 cells (results as for the wrapped input)."""
    ff = ForceField.bundled("ffield.reax.HO.2015")
    rc = ff.general[12]
    s, x = cases.water_cases()["distorted_dimer"]
    cell = np.diag([12.0, 13.0, 14.0])
    calc = Calculator(ff, qeq_tolerance=1e-12)
    expected = calc.evaluate(s, x, cell=cell, pbc=True)
    base = canonical_neighbors(orc.build_neighbors(x, rc, cell, True), len(x), cell, True)
    i0 = np.repeat(np.arange(len(x)), np.diff(base[0]))
    w = np.array([1000, 0, -700])
    x_far = x.copy()
    x_far[4] += w @ cell
    S_far = base[2] - np.where(base[1][:, None] == 4, w, 0) + np.where(i0[:, None] == 4, w, 0)
    supplied = (np.r_[i0, [0, 1, 2, 2]], np.r_[base[1], [1, 0, 2, 2]],
                np.vstack((S_far, [[2**40, 0, 0], [-(2**40), 0, 0], [0, 2**62 - 5, 0], [0, -(2**62 - 5), 0]])))
    print("[supplied] huge shifts", flush=True)
    got = calc.evaluate(s, x_far, cell=cell, pbc=True, neighbors=supplied)
    print("[built] far atoms", flush=True)
    built = calc.evaluate(s, x_far, cell=cell, pbc=True)
    for r in (got, built):
        assert abs(r.energy - expected.energy) <= 1e-9 * abs(expected.energy)
        np.testing.assert_allclose(r.forces, expected.forces, rtol=0, atol=1e-8)
        np.testing.assert_allclose(r.charges, expected.charges, rtol=0, atol=1e-9)
        np.testing.assert_allclose(r.bond_orders, expected.bond_orders, rtol=0, atol=1e-9)


@pytest.mark.gpu
def test_malformed_row_pointers_rejected():
    """A supplied CSR with bad row pointers fails: on the host in Engine.evaluate, and on the
    device when it reaches the C ABI directly (pointers clamped, the system flagged)."""
    import ctypes as C

    import torch

    from xreac_amd import _lib
    from xreac_amd.engine import Engine

    ff = ForceField.bundled("ffield.reax.HO.2015")
    eng = Engine(ff)
    s, x = cases.water_cases()["distorted_dimer"]
    row_ptr, j, S = canonical_neighbors(orc.build_neighbors(x, ff.general[12]), len(x))
    xd = torch.as_tensor(x, device=eng.device)
    types = torch.as_tensor(eng.type_ids(s), device=eng.device)
    bad = row_ptr.copy()
    bad[2], bad[3] = bad[3], bad[2]  # one decreasing step, last entry still the edge count
    for ptr in (bad, row_ptr - 1, np.r_[row_ptr[:-1], row_ptr[-1] + 1]):
        with pytest.raises(ValueError, match="row pointers"):
            eng.evaluate(xd, types, neighbors=(ptr, j, S))
    good = eng.evaluate(xd, types, neighbors=(row_ptr, j, S))
    # straight through the C ABI, bypassing the host check
    dev = [torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=eng.device) for a in (bad, j, S)]
    out = eng.outputs(len(s), properties=False)
    sysd = _lib.System(len(s), C.c_void_p(xd.data_ptr()), C.c_void_p(types.data_ptr()), None,
                       (C.c_int32 * 3)(0, 0, 0), 0.0, C.c_void_p(dev[0].data_ptr()), C.c_void_p(dev[1].data_ptr()),
                       C.c_void_p(dev[2].data_ptr()), len(j))
    opt = _lib.Options(1e-10, 2000, 1, 0, 0, 0, 0)
    res = _lib.Result()
    res.forces = C.c_void_p(out["forces"].data_ptr())
    res.charges = C.c_void_p(out["charges"].data_ptr())
    print("[supplied] decreasing row pointer through the C ABI", flush=True)
    rc = eng.lib.xreac_evaluate(eng._ctx, eng._ff, C.byref(sysd), C.byref(opt), C.byref(res))
    assert rc == _lib.XREAC_INVALID
    assert "malformed" in eng.lib.xreac_error_message(eng._ctx).decode()
    again = eng.evaluate(xd, types, neighbors=(row_ptr, j, S))  # the context is still usable
    np.testing.assert_array_equal(again["forces"].cpu().numpy(), good["forces"].cpu().numpy())
