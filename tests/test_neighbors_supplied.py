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
from xreac_amd.geometry import canonical_neighbors

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
    with pytest.raises(ValueError, match="511"):
        canonical_neighbors((i, j, shifts * 200), 2, np.eye(3), True)


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
