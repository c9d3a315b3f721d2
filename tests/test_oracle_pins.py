"""Pins for the CPU oracle: reference golden values, known answers, FD and invariants.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand, a wrong mixing rule) fails:
golden energies pin the forward sum of all terms, finite differences pin every
hand-derived derivative rule, the independent QEq assembly pins the charges,
and brute-force enumeration pins the neighbor set.
"""

from dataclasses import replace
from itertools import product

import numpy as np
import pytest
import torch

from oracle import geometry as og
from oracle import reaxff as orc
from tests import cases
from xreac_amd.forcefield import ForceField


@pytest.fixture(scope="module")
def water_ff():
    return ForceField.bundled("ffield.reax.HO.2015")


@pytest.fixture(scope="module")
def zno_ff():
    return ForceField.bundled("ffield.reax.ZnOH.2010")


@pytest.fixture(scope="module")
def cho_ff():
    return ForceField.bundled("ffield.reax.CHO.2008")


# ------------------------------------------------------------ golden values
def test_release_monomer_golden(water_ff):
    # .github/workflows/release.yml:75-82 of the reference
    r = orc.evaluate(water_ff, ["O", "H", "H"], [[0, 0, 0], [0.97, 0, 0], [-0.243, 0.94, 0]])
    assert abs(r["energy"] - (-248.34420920633704)) < 1e-8
    assert np.isfinite(r["forces"]).all()
    assert abs(r["charges"].sum()) < 1e-12
    ion = orc.evaluate(water_ff, ["O", "H"], [[0, 0, 0], [0.97, 0, 0]], total_charge=-1)
    assert abs(ion["charges"].sum() + 1) < 1e-12


def test_small_water_hydrogen_bond_golden(water_ff):
    # tests/test_reference.py:98-99, tests/test_small_cells.py:62 of the reference
    _, s, x, cell, pbc = cases.small_cell_cases()["water_4A"]
    r = orc.evaluate(water_ff, s, x, cell=cell, pbc=pbc)
    assert r["components"]["hydrogen_bond"] == pytest.approx(-0.3700520183, abs=1e-9)


# ----------------------------------------------------------- parser known answers
def test_parameter_reader_known_answers(zno_ff, tmp_path):
    # tests/test_calculator.py:207-216 of the reference
    assert zno_ff.checksum == "b5af5a65573ce60f32405a973e33f7cc89a90bc69e624a8ca89811eb67b1a07b"
    assert zno_ff.atoms["O"]["eta"] == 2 * 8.3122
    assert zno_ff.pairs["Zn", "O"]["r_vdW"] == 2 * 2.1414
    assert zno_ff.pairs["Zn", "O"]["r_p"] == pytest.approx((1.0548 - 1.6836) / 2)
    broken = tmp_path / "broken.ff"
    broken.write_text(zno_ff.path.read_text()[:500])
    with pytest.raises(ValueError):
        ForceField.from_file(broken)


def test_bundled_checksums():
    expected = {
        "ffield.reax.HO.2015": "526dea2d55d748caed94734139dda66064e194f74f46eee9b0f511eead38b7d1",
        "ffield.reax.CHO.2008": "c5240c07e664dde5626fcdb740be1bdbe5e7e34d919902603db0bf962fea9aee",
        "ffield.reax.TiOH.Kim2013": "b1568da9a894c1f12bc419bf968b1450251e86a20690a8177f1554928006721d",
    }
    for name, digest in expected.items():
        assert ForceField.bundled(name).checksum == digest


def test_wildcard_torsions_and_precedence(water_ff, tmp_path):
    # tests/test_general.py:21-48 of the reference
    assert water_ff.elements == ("H", "O", "X")
    assert water_ff.pairs["X", "O"]["gamma"] > 0
    assert water_ff.pairs["X", "O"]["De_s"] == 0
    np.testing.assert_array_equal(water_ff.torsions["H", "O", "O", "H"], [2.5, -4.0, 0.9, -2.5, -1.0])
    np.testing.assert_array_equal(water_ff.torsions["X", "O", "O", "X"], [0.5511, 25.415, 1.133, -5.1903, -1.0])
    np.testing.assert_array_equal(water_ff.torsions["O", "H", "O", "H"], [0, 0.1, 0.02, -2.5415, 0])
    lines = water_ff.path.read_text().splitlines()
    start = next(i for i, line in enumerate(lines) if "Nr of torsions" in line) + 1
    lines[start : start + 6] = lines[start + 3 : start + 6] + lines[start : start + 3]
    path = tmp_path / "reordered.ff"
    path.write_text("\n".join(lines) + "\n")
    reordered = ForceField.from_file(path)
    assert water_ff.torsions.keys() == reordered.torsions.keys()
    for key in water_ff.torsions:
        np.testing.assert_array_equal(water_ff.torsions[key], reordered.torsions[key])


def test_standard_format_variants(water_ff, tmp_path):
    # tests/test_general.py:51-66 of the reference
    text = water_ff.path.read_text().replace("50.0000", "5.00000D+1", 1)
    text = text[: text.index("  1    ! Nr of hydrogen bonds")]
    f = tmp_path / "without_hb.ff"
    f.write_text(text)
    parsed = ForceField.from_file(f)
    np.testing.assert_array_equal(parsed.general, water_ff.general)
    assert parsed.hydrogen_bonds == {}
    assert parsed.atoms.keys() == water_ff.atoms.keys()
    f.write_text("")
    with pytest.raises(ValueError, match="Empty"):
        ForceField.from_file(f)
    with pytest.raises(ValueError, match="plain filename"):
        ForceField.bundled("../pyproject.toml")


def test_lossless_export(cho_ff, tmp_path):
    out = tmp_path / "copy.ff"
    cho_ff.to_file(out)
    again = ForceField.from_file(out)
    np.testing.assert_array_equal(again.general, cho_ff.general)
    assert again.atoms == cho_ff.atoms
    assert again.section_counts == cho_ff.section_counts
    for key in cho_ff.torsions:
        np.testing.assert_array_equal(again.torsions[key], cho_ff.torsions[key])


# -------------------------------------------------------------- primitives
def test_scatter_sum_known_answer():
    # tests/test_neighbors.py:129-134 of the reference
    idx = torch.tensor([0, 1, 0, 3])
    x = torch.tensor([1.0, 2.0, 3.0, 4.0], dtype=torch.float64, requires_grad=True)
    y = orc.scatter_sum(x, idx, 5)
    np.testing.assert_array_equal(y.detach().numpy(), [4, 2, 0, 4, 0])
    (g,) = torch.autograd.grad(torch.sum(y**2), x)
    np.testing.assert_array_equal(g.numpy(), [8, 4, 8, 8])


def test_shift_offsets_match_numpy_matmul():
    # The reference forms S @ cell with NumPy (neighbors.py:44,125); the oracle
    # states that rounding explicitly.  Pin it on triclinic cells.
    rng = np.random.default_rng(5)
    for _ in range(5):
        lattice = rng.normal(size=(3, 3)) * 4
        S = rng.integers(-4, 5, size=(3000, 3))
        np.testing.assert_array_equal(og.shift_offsets(S, lattice), S @ lattice)


# -------------------------------------------------------------- neighbor sets
def test_neighbor_shift_completeness():
    # tests/test_neighbors.py:107-126 of the reference (brute-force images)
    cell = np.array([[3.1, 0.0, 0.0], [0.8, 3.6, 0.0], [0.2, 0.3, 4.1]])
    x = np.array([[0.0, 0.0, 0.0], [0.9, 0.2, 0.3]])
    pbc = np.array([True, True, False])
    i, j, S = og.image_neighbors(x, 7.5, cell, pbc)
    actual = {(int(a), int(b), *map(int, s)) for a, b, s in zip(i, j, S)}
    expected = og.brute_force_neighbors(x, 7.5, cell, pbc, 4)
    assert actual == expected
    assert max(abs(t[2]) for t in actual) >= 2
    order, reverse = og.canonical_edges(i, j, S, 2)
    np.testing.assert_array_equal(order, np.arange(len(i)))
    np.testing.assert_array_equal(reverse[reverse], np.arange(len(i)))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_neighbors_vs_brute_force_random(seed):
    rng = np.random.default_rng(seed)
    lattice = np.diag(rng.uniform(5, 9, 3)) + np.triu(rng.uniform(-1, 1, (3, 3)), 1)
    pbc = rng.integers(0, 2, 3).astype(bool)
    pbc[0] = True
    x = rng.uniform(-3, 12, size=(6, 3))  # unwrapped on purpose
    i, j, S = og.image_neighbors(x, 6.0, lattice, pbc)
    actual = {(int(a), int(b), *map(int, s)) for a, b, s in zip(i, j, S)}
    expected = og.brute_force_neighbors(x, 6.0, lattice, pbc, 5)
    assert actual == expected


def test_small_cell_self_images():
    # tests/test_small_cells.py:28-31: 80 self images within the 10 A cutoff
    _, s, x, cell, pbc = cases.small_cell_cases()["water_4A"]
    i, j, S = og.image_neighbors(x, 10.0, cell, np.array(pbc))
    self_counts = np.bincount(i[i == j], minlength=3)
    assert self_counts.tolist() == [80, 80, 80]
    assert np.abs(S).max() >= 2


# ------------------------------------------------------- finite differences
def _fd_check(ff, symbols, x, cell=None, pbc=None, total_charge=0, seed=21, h=1e-5, full=False):
    r = orc.evaluate(ff, symbols, x, cell=cell, pbc=pbc, total_charge=total_charge, full_derivative=full)
    x = np.asarray(x, dtype=float)
    direction = np.random.default_rng(seed).normal(size=x.shape)
    direction /= np.linalg.norm(direction)
    if full:
        plus = orc.evaluate(ff, symbols, x + h * direction, cell=cell, pbc=pbc, total_charge=total_charge)["energy"]
        minus = orc.evaluate(ff, symbols, x - h * direction, cell=cell, pbc=pbc, total_charge=total_charge)["energy"]
    else:
        model = r["model"]
        plus = orc.fixed_charge_energy(model, x + h * direction, r["charges"])
        minus = orc.fixed_charge_energy(model, x - h * direction, r["charges"])
    fd = (plus - minus) / (2 * h)
    assert fd == pytest.approx(-np.sum(r["forces"] * direction), abs=2e-5, rel=1e-6)
    np.testing.assert_allclose(r["forces"].sum(axis=0), 0, atol=1e-9)
    return r


def test_fd_water_dimer(water_ff):
    s, x = cases.water_cases()["distorted_dimer"]
    r = _fd_check(water_ff, s, x)
    assert r["components"]["hydrogen_bond"] < -1
    np.testing.assert_allclose(r["total_bond_orders"], r["bond_orders"].sum(axis=1), atol=1e-14)
    np.testing.assert_allclose(r["bond_orders"], r["bond_orders"].T, atol=1e-14)
    np.testing.assert_allclose(r["dipole"], (x * r["charges"][:, None]).sum(axis=0), atol=1e-12)


@pytest.mark.parametrize("name", ["zno", "o4", "cluster20"])
def test_fd_zno(zno_ff, name):
    s, x = cases.ZNO[name]
    _fd_check(zno_ff, s, x)


@pytest.mark.parametrize("name", ["zno", "o4"])
def test_fd_charge_response(zno_ff, name):
    s, x = cases.ZNO[name]
    _fd_check(zno_ff, s, x, full=True, seed=64)


@pytest.mark.parametrize("name", ["methane", "carbon_monoxide", "carbon_dimer"])
def test_fd_carbon(cho_ff, name):
    s, x = cases.carbon_cases()[name]
    r = _fd_check(cho_ff, s, x, h=1e-6, seed=508)
    if name == "carbon_dimer":
        assert r["components"]["lone_pair"] > 50  # C2 correction active


def test_fd_periodic_carbon_chain(cho_ff):
    _, s, x, cell, pbc = cases.periodic_carbon_chain()
    r = _fd_check(cho_ff, s, x, cell=cell, pbc=pbc, seed=192)
    assert abs(r["components"]["torsion"]) > 0.1


@pytest.mark.parametrize("name", ["water_4A", "partial_pbc_water"])
def test_fd_small_cells(water_ff, name):
    _, s, x, cell, pbc = cases.small_cell_cases()[name]
    _fd_check(water_ff, s, x, cell=cell, pbc=pbc, seed=130)


def test_fd_partial_pbc_slab(water_ff):
    s, x, cell, pbc = cases.periodic_cases()["partial_pbc_water"]
    _fd_check(water_ff, s, x, cell=cell, pbc=pbc, seed=900)


# ---------------------------------------------------------------- invariants
def test_symmetries_cluster20(zno_ff):
    s, x = cases.ZNO["cluster20"]
    rot, _ = np.linalg.qr(np.random.default_rng(33).normal(size=(3, 3)))
    r = orc.evaluate(zno_ff, s, x)
    moved = orc.evaluate(zno_ff, s, x @ rot + [4.2, -3.1, 7])
    assert moved["energy"] == pytest.approx(r["energy"], abs=1e-10)
    np.testing.assert_allclose(moved["forces"], r["forces"] @ rot, atol=1e-9)
    order = np.random.default_rng(33).permutation(len(x))
    perm = orc.evaluate(zno_ff, [s[k] for k in order], x[order])
    assert perm["energy"] == pytest.approx(r["energy"], abs=1e-10)
    np.testing.assert_allclose(perm["charges"], r["charges"][order], atol=1e-12)
    np.testing.assert_allclose(perm["forces"], r["forces"][order], atol=1e-9)


def test_periodic_wrapping(water_ff):
    s, x, cell, pbc = cases.periodic_cases()["partial_pbc_water"]
    r = orc.evaluate(water_ff, s, x, cell=cell, pbc=pbc)
    shifts = np.random.default_rng(900).integers(-3, 4, x.shape) * np.array(pbc)
    w = orc.evaluate(water_ff, s, x + shifts @ cell, cell=cell, pbc=pbc)
    assert w["energy"] == pytest.approx(r["energy"], abs=1e-9)
    np.testing.assert_allclose(w["forces"], r["forces"], atol=2e-8)
    np.testing.assert_allclose(w["charges"], r["charges"], atol=1e-11)


def test_supercell_extensivity(water_ff):
    s, x, cell, pbc = cases.periodic_cases()["boundary_dimer"]
    r = orc.evaluate(water_ff, s, x, cell=cell)
    big = cell.copy()
    big[0] *= 2
    d = orc.evaluate(water_ff, s * 2, np.vstack((x, x + cell[0])), cell=big)
    assert d["energy"] == pytest.approx(2 * r["energy"], abs=1e-9)
    for k in ("charges", "forces", "total_bond_orders", "lone_pairs", "bond_counts"):
        np.testing.assert_allclose(d[k], np.concatenate([r[k]] * 2), atol=1e-9)


def test_zinc_chain_self_image_bonds(zno_ff):
    # tests/test_reference.py:102-106: one atom bonded to its own images
    _, s, x, cell, pbc = cases.small_cell_cases()["zinc_chain"]
    r = orc.evaluate(zno_ff, s, x, cell=cell, pbc=pbc)
    assert r["bond_counts"].tolist() == [2]
    assert r["bond_orders"][0, 0] > 1.0
    assert r["total_bond_orders"][0] == r["bond_orders"][0, 0]
    np.testing.assert_allclose(r["forces"], 0, atol=1e-12)


def test_fixed_vs_charge_response_gap(zno_ff):
    s, x = cases.ZNO["zno"]
    fixed = orc.evaluate(zno_ff, s, x)
    full = orc.evaluate(zno_ff, s, x, full_derivative=True)
    assert fixed["energy"] == full["energy"]
    assert np.max(abs(full["forces"] - fixed["forces"])) > 0.03


def test_collinear_active_torsion_raises(cho_ff):
    s, x = ["C"] * 4, [[0, 0, 0], [1.5, 0, 0], [3, 0, 0], [4.5, 0.3, 0]]
    with pytest.raises(ValueError, match="Collinear atoms in an active torsion"):
        orc.evaluate(cho_ff, s, x)
    torsions = {k: v.copy() for k, v in cho_ff.torsions.items()}
    for v in torsions.values():
        v[[0, 1, 2, 4]] = 0
    inactive = replace(cho_ff, torsions=torsions)
    r = orc.evaluate(inactive, s, x)
    assert r["components"]["torsion"] == r["components"]["conjugation"] == 0


def test_coincident_atoms_rejected(water_ff):
    with pytest.raises(ValueError, match="Coincident"):
        orc.evaluate(water_ff, ["O", "H"], [[0.0, 0, 0], [12.0, 0, 0]], cell=[12.0] * 3)


# ----------------------------------------------------- independent QEq checks
def _independent_charges(ff, symbols, x, lattice, pbc, total_charge):
    """Constraint-elimination QEq from explicit images (scripts/validate.py:172-209)."""
    cutoff = ff.general[12]
    n = len(symbols)
    extent = 4 if np.any(pbc) else 0
    pairs = og.brute_force_neighbors(x, cutoff, lattice, pbc, extent)
    hess = np.diag([ff.atoms[s]["eta"] for s in symbols])
    for a, b, *s in pairs:
        v = x[b] - x[a] + np.array(s) @ lattice
        d = np.linalg.norm(v)
        if d >= cutoff:
            continue
        u = d / cutoff
        taper = 1 - 35 * u**4 + 84 * u**5 - 70 * u**6 + 20 * u**7
        gamma = (ff.atoms[symbols[a]]["gamma"] * ff.atoms[symbols[b]]["gamma"]) ** -1.5
        hess[a, b] += 14.4 * taper / (d**3 + gamma) ** (1 / 3)
    chi = np.array([ff.atoms[s]["chi"] for s in symbols])
    part = np.zeros(n)
    part[-1] = total_charge
    basis = np.vstack((np.eye(n - 1), -np.ones(n - 1)))
    reduced = basis.T @ hess @ basis
    return part + basis @ np.linalg.solve(reduced, -basis.T @ (hess @ part + chi))


@pytest.mark.parametrize("name", ["hydroxide", "hydronium", "charged_water_4A", "charged_partial_pbc_water"])
def test_independent_qeq(water_ff, name):
    s, x, cell, pbc, q = cases.charged_cases()[name]
    r = orc.evaluate(water_ff, s, x, cell=cell, pbc=pbc, total_charge=q)
    lattice = np.asarray(cell, dtype=float) if cell is not None else np.eye(3)
    flags = np.array(pbc if pbc is not False else [False] * 3, dtype=bool)
    expected = _independent_charges(water_ff, s, np.asarray(x), lattice, flags, q)
    np.testing.assert_allclose(r["charges"], expected, atol=1e-10)
    assert abs(r["charges"].sum() - q) < 1e-12


# ------------------------------------------------------------ integrators
def test_md_unit_constant_against_si():
    """A/fs^2 per kcal/mol/A/amu, checked through the 300 K velocity scale of oxygen in SI units."""
    from oracle import dynamics as od

    kb_si, amu, kb_kcal = 1.380649e-23, 1.66053906660e-27, 0.0019872043
    sigma_si = np.sqrt(kb_si * 300.0 / (15.999 * amu)) * 1e10 / 1e15  # A/fs
    sigma = np.sqrt(kb_kcal * 300.0 * od.ACCEL / 15.999)
    assert sigma == pytest.approx(sigma_si, rel=1e-6)


def test_velocity_verlet_harmonic_oscillator():
    from oracle import dynamics as od

    k, m, x0 = 2.0, 4.0, 0.3
    omega = np.sqrt(k * od.ACCEL / m)  # rad/fs

    def harmonic(x):
        return 0.5 * k * np.sum(x * x), -k * x

    x = np.array([[x0, 0.0, 0.0]])
    dt, steps = 0.05, 2000
    xt, vt, pot, kin = od.velocity_verlet(harmonic, x, np.zeros_like(x), [m], dt, steps)
    t = dt * steps
    # velocity Verlet: phase error O(dt^2 omega^3 t), energy bounded, no drift
    assert xt[0, 0] == pytest.approx(x0 * np.cos(omega * t), abs=1e-4)
    assert vt[0, 0] == pytest.approx(-x0 * omega * np.sin(omega * t), abs=1e-6)
    total = pot + kin
    assert np.ptp(total) < 1e-4 * total[0]


def test_fire_first_iterations_closed_form():
    """Hand-derived first FIRE moves on E = k|x|^2/2 (calculator.py:225-245)."""
    from oracle import dynamics as od

    x0 = np.array([[1.0, 0.0, 0.0]])
    # k = 1: it0 v = 0.02 F0, x1 = x0 + 0.0004 F0; it1 power > 0 and v is parallel to F,
    # so the mixing keeps v, then v += 0.02 F1 and x2 = x1 + 0.0004 (F0 + F1).
    _, _, _, traj = od.fire(lambda x: -x, x0, 1e-12, 2)
    assert traj[1][0, 0] == pytest.approx(0.9996, abs=1e-15)
    assert traj[2][0, 0] == pytest.approx(0.9996 - 0.0004 * 1.9996, abs=1e-15)
    # k = 1000: the first step 0.4 A is capped to 0.1 A
    _, _, _, traj = od.fire(lambda x: -1000.0 * x, x0, 1e-12, 1)
    assert traj[1][0, 0] == pytest.approx(0.9, abs=1e-15)
    # convergence is tested before moving
    x, converged, it, traj = od.fire(lambda x: -x * 1e-9, x0, 1e-6, 5)
    assert converged and it == 0 and len(traj) == 1


def test_oracle_full_derivative_is_the_energy_gradient(zno_ff):
    """Charge-response forces of the oracle against central differences of the equilibrated energy on the
    reference's ZnO dimer (tests/test_calculator.py:109-118: gap to fixed-charge forces > 0.03)."""
    s, x = cases.ZNO["zno"]
    x = np.asarray(x, dtype=np.float64)
    full = orc.evaluate(zno_ff, s, x, full_derivative=True)
    fixed = orc.evaluate(zno_ff, s, x)
    assert np.max(np.abs(full["forces"] - fixed["forces"])) > 0.03
    h = 1e-5
    for a, k in ((0, 0), (1, 0), (1, 1)):
        xp, xm = x.copy(), x.copy()
        xp[a, k] += h
        xm[a, k] -= h
        fd = -(orc.evaluate(zno_ff, s, xp)["energy"] - orc.evaluate(zno_ff, s, xm)["energy"]) / (2 * h)
        assert full["forces"][a, k] == pytest.approx(fd, abs=1e-6)
