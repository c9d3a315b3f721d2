"""GPU path vs the CPU oracle on the same inputs (run on an MI355X).

Bars (from the path's specification):
  neighbor list and active-bond set  - bit-exact, canonical order
  energy                             - 1e-9 relative
  forces                             - 1e-6 kcal/mol/A max-abs
  charges                            - 1e-7 e at QEq tolerance 1e-10 (tests use 1e-12)
  per-atom bond properties           - 1e-8; bond counts exact
"""

import numpy as np
import pytest

from oracle import reaxff as orc
from tests import cases
from xreac_amd import Calculator, ForceField, workloads

pytestmark = pytest.mark.gpu

E_REL = 1e-9
F_ABS = 1e-6
Q_ABS = 1e-7


def _systems():
    out = {}
    for name, (s, x) in cases.water_cases().items():
        out["water_" + name] = ("ffield.reax.HO.2015", s, x, None, None, 0.0)
    for name, (s, x) in cases.ZNO.items():
        out["zno_" + name] = ("ffield.reax.ZnOH.2010", s, x, None, None, 0.0)
    for name, (s, x) in cases.carbon_cases().items():
        out["cho_" + name] = ("ffield.reax.CHO.2008", s, x, None, None, 0.0)
    out["periodic_carbon_chain"] = (*cases.periodic_carbon_chain(), 0.0)
    for name, (ffn, s, x, cell, pbc) in cases.small_cell_cases().items():
        out["small_" + name] = (ffn, s, x, cell, pbc, 0.0)
    for name, (s, x, cell, pbc) in cases.periodic_cases(include_bulk=True).items():
        out["periodic_" + name] = ("ffield.reax.HO.2015", s, x, cell, pbc, 0.0)
    for name, (s, x, cell, pbc, q) in cases.charged_cases().items():
        out["charged_" + name] = ("ffield.reax.HO.2015", s, x, cell, pbc, q)
    s, x = workloads.water_cluster()
    out["water_cluster_192"] = ("ffield.reax.HO.2015", s, x, None, None, 0.0)
    s, x, cell = workloads.cho_mixture(counts=(("CH4", 12), ("O2", 12), ("H2O", 6), ("CH3OH", 4)), length=14.0, seed=7)
    out["cho_mixture_small"] = ("ffield.reax.CHO.2008", s, x, cell, [True] * 3, 0.0)
    s, x = workloads.anatase_cell()
    x = x + np.random.default_rng(4).normal(0, 0.02, x.shape)
    out["tioh_anatase_bulk"] = ("ffield.reax.TiOH.Kim2013", s, x,
                                np.diag([workloads.ANATASE_A, workloads.ANATASE_A, workloads.ANATASE_C]), [True] * 3, 0.0)
    s, x, cell = workloads.anatase_water_slab(nx=2, ny=2, nz=1, waters=8, water_thickness=8.0, vacuum=6.0, seed=5)
    out["tioh_slab_water"] = ("ffield.reax.TiOH.Kim2013", s, x, cell, [True] * 3, 0.0)
    # ~610 neighbors per atom within 10 A: exercises rows longer than 512
    rng = np.random.default_rng(3)
    grid = np.array(np.meshgrid(*[np.arange(10)] * 3, indexing="ij")).reshape(3, -1).T
    x = grid * 1.9 + rng.normal(0, 0.1, grid.shape)
    out["dense_oxygen_lattice"] = ("ffield.reax.HO.2015", ["O"] * len(x), x, np.diag([19.0] * 3), [True] * 3, 0.0)
    return out


SYSTEMS = _systems()
_CALCS = {}


def _calc(ffname):
    if ffname not in _CALCS:
        _CALCS[ffname] = Calculator(ForceField.bundled(ffname), qeq_tolerance=1e-12, max_expanded_atoms=10**9)
    return _CALCS[ffname]


def assert_matches_oracle(calc, s, x, cell=None, pbc=None, q=0.0, full_derivative=False):
    """One evaluation on the device against the oracle at every bar of the module docstring."""
    gpu = calc.evaluate(s, x, cell=cell, pbc=pbc, total_charge=q, full_derivative=full_derivative)
    ref = orc.evaluate(calc.force_field, s, x, cell=cell, pbc=pbc, total_charge=q, full_derivative=full_derivative)
    model = ref["model"]
    # neighbor list: identical canonical (i, j, S)
    eng = calc.engine
    m = len(model.i)
    row_ptr, j, S = eng.neighbors(len(s), m)
    np.testing.assert_array_equal(row_ptr, model.edges.row_ptr)
    np.testing.assert_array_equal(j, model.j)
    np.testing.assert_array_equal(S, model.edges.shifts)
    # active bonds: the candidate edges whose raw bond order passes the cutoff
    active = model._bond_candidates[model._bond_mask]
    bi, bj, bS, bo = eng.bonds(len(active))
    np.testing.assert_array_equal(bi, model.i[active])
    np.testing.assert_array_equal(bj, model.j[active])
    np.testing.assert_array_equal(bS, model.edges.shifts[active])
    e_ref = ref["energy"]
    assert abs(gpu.energy - e_ref) <= E_REL * max(1.0, abs(e_ref)), (gpu.energy, e_ref)
    for k, v in ref["components"].items():
        assert abs(gpu.components[k] - v) <= E_REL * max(1.0, abs(e_ref)), (k, gpu.components[k], v)
    np.testing.assert_allclose(gpu.forces, ref["forces"], rtol=0, atol=F_ABS)
    np.testing.assert_allclose(gpu.charges, ref["charges"], rtol=0, atol=Q_ABS)
    np.testing.assert_allclose(gpu.total_bond_orders, ref["total_bond_orders"], rtol=0, atol=1e-8)
    np.testing.assert_allclose(gpu.lone_pairs, ref["lone_pairs"], rtol=0, atol=1e-8)
    np.testing.assert_array_equal(gpu.bond_counts, ref["bond_counts"])
    np.testing.assert_allclose(gpu.bond_orders, ref["bond_orders"], rtol=0, atol=1e-8)
    np.testing.assert_allclose(gpu.dipole, ref["dipole"], rtol=0, atol=1e-7)
    assert abs(gpu.charges.sum() - q) < 1e-10
    return gpu, ref


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_parity(name):
    ffname, s, x, cell, pbc, q = SYSTEMS[name]
    print(f"[parity] {name}: n={len(s)}", flush=True)
    assert_matches_oracle(_calc(ffname), s, x, cell, pbc, q)


def test_golden_monomer_on_gpu():
    calc = _calc("ffield.reax.HO.2015")
    r = calc.evaluate(["O", "H", "H"], [[0, 0, 0], [0.97, 0, 0], [-0.243, 0.94, 0]])
    assert abs(r.energy - (-248.34420920633704)) < 1e-8
    assert abs(r.charges.sum()) < 1e-12
    ion = calc.evaluate(["O", "H"], [[0, 0, 0], [0.97, 0, 0]], total_charge=-1)
    assert abs(ion.charges.sum() + 1) < 1e-12


def test_golden_small_water_hbond_on_gpu():
    _, s, x, cell, pbc = cases.small_cell_cases()["water_4A"]
    r = _calc("ffield.reax.HO.2015").evaluate(s, x, cell=cell, pbc=pbc)
    assert r.components["hydrogen_bond"] == pytest.approx(-0.3700520183, abs=1e-9)


def test_errors_on_gpu():
    water = _calc("ffield.reax.HO.2015")
    with pytest.raises(ValueError, match="Coincident"):
        water.evaluate(["O", "H"], [[0.0, 0, 0], [12.0, 0, 0]], cell=[12.0] * 3)
    with pytest.raises(ValueError, match="absent from the force field"):
        water.evaluate(["Zn"], [[0, 0, 0]])
    cho = _calc("ffield.reax.CHO.2008")
    with pytest.raises(ValueError, match="Collinear atoms in an active torsion"):
        cho.evaluate(["C"] * 4, [[0, 0, 0], [1.5, 0, 0], [3, 0, 0], [4.5, 0.3, 0]])


def test_qeq_nonconvergence_reported():
    import torch

    calc = _calc("ffield.reax.HO.2015")
    s, x = workloads.water_cluster()
    eng = calc.engine
    with pytest.raises(ValueError, match="Singular QEq system.*did not converge"):
        eng.evaluate(torch.as_tensor(x, device=eng.device), torch.as_tensor(eng.type_ids(s), device=eng.device),
                     max_iter=1)
    ok = eng.evaluate(torch.as_tensor(x, device=eng.device), torch.as_tensor(eng.type_ids(s), device=eng.device))
    assert ok["qeq_iterations"] > 1


def test_buffer_hints_grow_and_match_fresh_context():
    """A context sized by a small system must still give exact results for a denser one."""
    import torch

    from xreac_amd.engine import Engine

    ff = ForceField.bundled("ffield.reax.HO.2015")
    used = Engine(ff)
    s1, x1 = cases.water_cases()["distorted_dimer"]
    used.evaluate(torch.as_tensor(x1, device=used.device), torch.as_tensor(used.type_ids(s1), device=used.device))
    s2, x2, cell2 = workloads.water_box(side=6, seed=4)
    args = (torch.as_tensor(x2, device=used.device), torch.as_tensor(used.type_ids(s2), device=used.device))
    grown = used.evaluate(*args, cell=cell2, pbc=(True,) * 3)
    fresh_engine = Engine(ff)
    fresh = fresh_engine.evaluate(*args, cell=cell2, pbc=(True,) * 3)
    assert grown["num_edges"] == fresh["num_edges"] and grown["num_bonds"] == fresh["num_bonds"]
    np.testing.assert_array_equal(grown["components"], fresh["components"])
    np.testing.assert_array_equal(grown["forces"].cpu().numpy(), fresh["forces"].cpu().numpy())
    again = used.evaluate(*args, cell=cell2, pbc=(True,) * 3)  # now from hints, no read-back
    np.testing.assert_array_equal(again["forces"].cpu().numpy(), fresh["forces"].cpu().numpy())


def test_deterministic_repeat():
    s, x = workloads.water_cluster()
    calc = _calc("ffield.reax.HO.2015")
    a = calc.evaluate(s, x)
    b = calc.evaluate(s, x)
    assert a.energy == b.energy
    np.testing.assert_array_equal(a.forces, b.forces)
    np.testing.assert_array_equal(a.charges, b.charges)


FULL = ["water_distorted_dimer", "water_cluster_192", "periodic_bulk_water_192", "charged_hydronium",
        "charged_charged_water_4A", "small_water_4A", "cho_mixture_small", "zno_cluster20", "periodic_carbon_chain"]


@pytest.mark.parametrize("name", FULL)
def test_parity_full_derivative(name):
    """Charge-response forces (calculator.py:79-86) against the oracle's autograd through the QEq solve."""
    ffname, s, x, cell, pbc, q = SYSTEMS[name]
    print(f"[full] {name}: n={len(s)}", flush=True)
    calc = _calc(ffname)
    gpu = calc.evaluate(s, x, cell=cell, pbc=pbc, total_charge=q, full_derivative=True)
    ref = orc.evaluate(calc.force_field, s, x, cell=cell, pbc=pbc, total_charge=q, full_derivative=True)
    assert gpu.full_derivative and gpu.force_convention == "charge_response"
    assert abs(gpu.energy - ref["energy"]) <= E_REL * max(1.0, abs(ref["energy"]))
    np.testing.assert_allclose(gpu.forces, ref["forces"], rtol=0, atol=F_ABS)
    np.testing.assert_allclose(gpu.charges, ref["charges"], rtol=0, atol=Q_ABS)


def test_full_derivative_is_the_energy_gradient():
    """Charge-response forces against central differences of the QEq-equilibrated energy (h = 1e-5 A),
    on the reference's ZnO dimer (tests/test_calculator.py:109-118)."""
    calc = _calc("ffield.reax.ZnOH.2010")
    s, x = cases.ZNO["zno"]
    x = np.asarray(x, dtype=np.float64)
    full = calc.evaluate(s, x, full_derivative=True)
    fixed = calc.evaluate(s, x)
    assert fixed.energy == full.energy
    np.testing.assert_array_equal(fixed.charges, full.charges)
    assert fixed.force_convention == "fixed_charge" and full.force_convention == "charge_response"
    assert np.max(np.abs(full.forces - fixed.forces)) > 0.03
    h = 1e-5
    fd = np.zeros_like(x)
    for a in range(len(s)):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, k] += h
            xm[a, k] -= h
            fd[a, k] = -(calc.evaluate(s, xp).energy - calc.evaluate(s, xm).energy) / (2 * h)
    np.testing.assert_allclose(full.forces, fd, rtol=0, atol=2e-5)
