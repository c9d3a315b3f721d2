"""Closed-form pins for the oracle's many-body and correction terms.

The reference holds no golden value for these terms, so each check here is one
the mathematics fixes.  Parameters are edited (``cases.modified``) so that a
term collapses to an expression of the geometry and the file's columns alone,
and then each dropped factor is restored and checked against the bond orders
that the public API returns:

* torsion and conjugation (energy.py:320-370): with p(tor2) -> inf and
  p(cot2) = 0, E_tor(omega) = (s1 s2 / 2) [V1 (1 + cos w) + V2 (1 - cos 2w)
  + V3 (1 + cos 3w)] and E_conj = cot1 (1 - s1 s2 sin^2 w), w the IUPAC
  dihedral (0 = cis), s1, s2 the sines of the two bond angles;
* angle penalty and angle conjugation (energy.py:310-318): pen1 and coa1 / 2
  per angle once every exponent is zero; invariant when the angle opens at
  fixed bond lengths; linear in the row's pen1 and coa1;
* triple-bond stabilisation (energy.py:239-246): p(trip1) / 13 per C-O pair
  with bo >= 1 once p(trip2) = p(trip3) = p(trip4) = 0, counted once per
  pair, zero below bo = 1;
* C2 correction (energy.py:233-237): linear in kc2, zero below vov3 = 3,
  summed over both directions of the C-C edge;
* lammps_angle (energy.py:49-55): above sin(theta) = 1e-5 the derivative is
  the finite difference of the energy; below it the theta-part of the
  gradient is scaled by sin(theta) / 1e-5; the plain arm clips at 1 - 1e-14
  and drops the theta-part (energy.py:304).

A swapped V1/V3 column, a sign flip in cos w, a wrong constant in any
bracket, a missing sin(theta) product, a doubled triple-bond count, a power
other than 2 in the C2 term or a clamp other than 1e-5 each fail one of them.
"""

import numpy as np
import pytest
import torch

from oracle import reaxff as orc
from tests import cases
from xreac_amd.forcefield import ForceField

OMEGAS = np.linspace(0.0, 2.0 * np.pi, 24, endpoint=False)
HCCH = ("H", "C", "C", "H")
# general-parameter indices (file order, 0-based)
P_TRIP4, P_TRIP3, KC2, P_TRIP2, P_TRIP1 = 3, 4, 5, 7, 10
P_COA2, P_PEN2, P_PEN3, P_PEN4, P_COA4, P_COA3 = 2, 19, 20, 21, 30, 38
P_TOR2, P_TOR3, P_COT2 = 23, 24, 27


@pytest.fixture(scope="module")
def cho():
    return ForceField.bundled("ffield.reax.CHO.2008")


@pytest.fixture(scope="module")
def ho():
    return ForceField.bundled("ffield.reax.HO.2015")


def _sweep(ff, omegas, lengths=(1.09, 1.54, 1.09), angles=(125.0, 135.0)):
    tor, conj, bos = [], [], []
    for w in omegas:
        x = cases.dihedral_chain(w, lengths, angles)
        assert cases.iupac_dihedral(x) == pytest.approx(np.arctan2(np.sin(w), np.cos(w)), abs=1e-12)
        r = orc.evaluate(ff, list(HCCH), x)
        assert r["model"].torsion_list.shape == (1, 3)  # one chain: no 1-3 or 1-4 contact is active
        tor.append(r["components"]["torsion"])
        conj.append(r["components"]["conjugation"])
        bos.append(r["bond_orders"])
    bos = np.array(bos)
    # the bond orders do not depend on the dihedral, so only the angular factor varies
    assert np.max(np.abs(bos - bos[0])) < 1e-14
    return np.array(tor), np.array(conj), bos[0]


# ------------------------------------------------------------------ torsions
@pytest.mark.parametrize("angles", [(125.0, 135.0), (140.0, 118.0)])
def test_torsion_fourier_closed_form(cho, angles):
    s12 = np.sin(np.radians(angles[0])) * np.sin(np.radians(angles[1]))
    w = OMEGAS
    for row in ([1.0, 0, 0, 0, 0], [0, 1.0, 0, 0, 0], [0, 0, 1.0, 0, 0], [-0.7, 3.1, 0.45, 0, 0]):
        ff = cases.modified(cho, general={P_TOR2: 1e3, P_COT2: 0.0}, torsions={HCCH: row})
        tor, conj, _ = _sweep(ff, w, angles=angles)
        v1, v2, v3 = row[:3]
        closed = 0.5 * s12 * (v1 * (1 + np.cos(w)) + v2 * (1 - np.cos(2 * w)) + v3 * (1 + np.cos(3 * w)))
        np.testing.assert_allclose(tor, closed, rtol=0, atol=1e-13)
        np.testing.assert_array_equal(conj, 0.0)
    # Fourier coefficients of the single-column rows: the harmonic each column multiplies and its sign
    basis = np.stack([np.ones_like(w), np.cos(w), np.cos(2 * w), np.cos(3 * w), np.sin(w), np.sin(2 * w)], axis=1)
    for col, harmonic, sign in ((0, 1, 1.0), (1, 2, -1.0), (2, 3, 1.0)):
        row = [0.0] * 5
        row[col] = 1.0
        ff = cases.modified(cho, general={P_TOR2: 1e3, P_COT2: 0.0}, torsions={HCCH: row})
        tor, _, _ = _sweep(ff, w, angles=angles)
        coef, *_ = np.linalg.lstsq(basis, tor, rcond=None)
        expect = np.zeros(6)
        expect[0] = 0.5 * s12
        expect[harmonic] = sign * 0.5 * s12
        np.testing.assert_allclose(coef, expect, rtol=0, atol=1e-13)


def test_conjugation_closed_form(cho):
    w = OMEGAS
    for angles in ((125.0, 135.0), (140.0, 118.0)):
        s12 = np.sin(np.radians(angles[0])) * np.sin(np.radians(angles[1]))
        for cot1 in (1.0, -2.2):
            ff = cases.modified(cho, general={P_TOR2: 1e3, P_COT2: 0.0}, torsions={HCCH: [0, 0, 0, 0, cot1]})
            tor, conj, _ = _sweep(ff, w, angles=angles)
            np.testing.assert_array_equal(tor, 0.0)
            # a + b cos 2w with b / a = s12 / (2 - s12)
            np.testing.assert_allclose(conj, cot1 * (1 - s12 * np.sin(w) ** 2), rtol=0, atol=1e-13)


def test_torsion_bond_order_factors(cho):
    """fn10 = prod(1 - exp(-p(tor2) b)) and fn12 = exp(-p(cot2) sum (b - 1.5)^2) over the
    three bonds, b = BO - 0.001 from the public bond-order matrix."""
    row = [1.3, 0.0, 0.4, 0.0, -1.7]
    w = OMEGAS[:6]
    limit = cases.modified(cho, general={P_TOR2: 1e3, P_COT2: 0.0}, torsions={HCCH: row})
    real = cases.modified(cho, torsions={HCCH: row})
    t_inf, c_0, bo = _sweep(limit, w)
    t_real, c_real, bo2 = _sweep(real, w)
    np.testing.assert_array_equal(bo, bo2)
    b = np.array([bo[0, 1], bo[1, 2], bo[2, 3]]) - 0.001
    fn10 = np.prod(1 - np.exp(-cho.general[P_TOR2] * b))
    fn12 = np.exp(-cho.general[P_COT2] * np.sum((b - 1.5) ** 2))
    np.testing.assert_allclose(t_real, fn10 * t_inf, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(c_real, fn12 * c_0, rtol=1e-13, atol=1e-15)


def test_torsion_tor1_factor(cho):
    """ex1 = exp(tor1 (2 - pi_jk - f11)^2): with pi bonds switched off on C and f11 -> 1
    (p(tor3) large, the C atoms under-coordinated), E_tor scales by exactly exp(tor1)."""
    base = dict(general={P_TOR2: 1e3, P_COT2: 0.0, P_TOR3: 50.0}, atoms={"C": {"r_pi": -1.0}})
    w = OMEGAS[1:7]
    ref = None
    for tor1 in (0.0, -0.5, -2.5):
        ff = cases.modified(cho, torsions={HCCH: [0, 1.0, 0, tor1, 0]}, **base)
        tor, _, bo = _sweep(ff, w)
        total = bo.sum(axis=1)
        assert np.all(total[1:3] < ff.atoms["C"]["valency_boc"])  # Delta_boc < 0 at both middle atoms
        if ref is None:
            ref = tor
            s12 = np.sin(np.radians(125.0)) * np.sin(np.radians(135.0))
            np.testing.assert_allclose(ref, 0.5 * s12 * (1 - np.cos(2 * w)), rtol=0, atol=1e-13)
        np.testing.assert_allclose(tor, np.exp(tor1) * ref, rtol=1e-13, atol=1e-15)


# ---------------------------------------------------- penalty and conjugation
OCO = ["O", "C", "O"]


def _triatomic(ff, theta_deg, lengths=(1.4, 1.4)):
    r = orc.evaluate(ff, OCO, cases.bent_triatomic(np.radians(theta_deg), lengths))
    assert len(r["model"].angle_list[0]) == 1  # the O-C-O angle only (no O-O contact)
    return r


def test_penalty_and_angle_conjugation_closed_form(cho):
    pen1, coa1 = cho.angles["O", "C", "O"][0][5], cho.angles["O", "C", "O"][0][3]
    assert pen1 != 0 and coa1 != 0
    flat = cases.modified(cho, general={P_PEN2: 0.0, P_PEN3: 0.0, P_PEN4: 0.0, P_COA2: 0.0, P_COA3: 0.0, P_COA4: 0.0})
    for theta in (120.0, 150.0):
        c = _triatomic(flat, theta)["components"]
        assert c["penalty"] == pytest.approx(pen1, rel=1e-15)
        assert c["angle_conjugation"] == pytest.approx(0.5 * coa1, rel=1e-15)


def test_penalty_and_angle_conjugation_invariant_in_theta(cho):
    out = [_triatomic(cho, t) for t in (120.0, 135.0, 150.0, 165.0)]
    for r in out[1:]:
        np.testing.assert_array_equal(r["bond_orders"], out[0]["bond_orders"])
        assert r["components"]["penalty"] == out[0]["components"]["penalty"]
        assert r["components"]["angle_conjugation"] == out[0]["components"]["angle_conjugation"]
    angle = [r["components"]["angle"] for r in out]
    assert np.all(np.abs(np.diff(angle)) > 1.0)  # the valence term itself does change


def test_penalty_and_angle_conjugation_linear_in_row(cho):
    row = cho.angles["O", "C", "O"][0]
    base = _triatomic(cho, 130.0)["components"]
    for scale in (0.0, 2.0, -0.5):
        edited = row.copy()
        edited[5] *= scale
        edited[3] *= scale
        c = _triatomic(cases.modified(cho, angles={("O", "C", "O"): [edited]}), 130.0)["components"]
        assert c["penalty"] == pytest.approx(scale * base["penalty"], rel=1e-14, abs=1e-300)
        assert c["angle_conjugation"] == pytest.approx(scale * base["angle_conjugation"], rel=1e-14, abs=1e-300)
        assert c["angle"] == base["angle"]


def test_penalty_and_angle_conjugation_factors(cho):
    """Each factor restored on its own, against the public bond orders:
    penalty  pen1 * (2 + e^(-p3 d)) / (1 + e^(-p3 d) + e^(p4 d)) * e^(-p2 ((b1-2)^2 + (b2-2)^2))
    coa      coa1 / (1 + e^(c2 (T_c - val_val))) * e^(-c3 ((T_1 - b1)^2 + (T_2 - b2)^2))
                  * e^(-c4 ((b1-1.5)^2 + (b2-1.5)^2))
    with d = T_c - valency of the centre, T the total bond orders, b = BO - 0.001."""
    g = cho.general
    row = cho.angles["O", "C", "O"][0]
    pen1, coa1 = row[5], row[3]
    zero = {P_PEN2: 0.0, P_PEN3: 0.0, P_PEN4: 0.0, P_COA2: 0.0, P_COA3: 0.0, P_COA4: 0.0}
    lengths = (1.33, 1.42)
    r = _triatomic(cho, 128.0, lengths)
    bo, tot = r["bond_orders"], r["total_bond_orders"]
    b1, b2 = bo[1, 0] - 0.001, bo[1, 2] - 0.001
    d = tot[1] - cho.atoms["C"]["valency"]

    def comp(keep):
        gen = {k: v for k, v in zero.items() if k not in keep}
        return _triatomic(cases.modified(cho, general=gen), 128.0, lengths)["components"]

    c = comp({P_PEN2})
    assert c["penalty"] == pytest.approx(pen1 * np.exp(-g[P_PEN2] * ((b1 - 2) ** 2 + (b2 - 2) ** 2)), rel=1e-13)
    c = comp({P_PEN3, P_PEN4})
    e3, e4 = np.exp(-g[P_PEN3] * d), np.exp(g[P_PEN4] * d)
    assert c["penalty"] == pytest.approx(pen1 * (2 + e3) / (1 + e3 + e4), rel=1e-13)
    c = comp({P_COA2})
    assert c["angle_conjugation"] == pytest.approx(
        coa1 / (1 + np.exp(g[P_COA2] * (tot[1] - cho.atoms["C"]["valency_val"]))), rel=1e-13)
    c = comp({P_COA3})
    assert c["angle_conjugation"] == pytest.approx(
        0.5 * coa1 * np.exp(-g[P_COA3] * ((tot[0] - b1) ** 2 + (tot[2] - b2) ** 2)), rel=1e-13)
    c = comp({P_COA4})
    assert c["angle_conjugation"] == pytest.approx(
        0.5 * coa1 * np.exp(-g[P_COA4] * ((b1 - 1.5) ** 2 + (b2 - 1.5) ** 2)), rel=1e-13)
    # all factors together reproduce the unedited term
    full = r["components"]
    e3, e4 = np.exp(-g[P_PEN3] * d), np.exp(g[P_PEN4] * d)
    assert full["penalty"] == pytest.approx(
        pen1 * (2 + e3) / (1 + e3 + e4) * np.exp(-g[P_PEN2] * ((b1 - 2) ** 2 + (b2 - 2) ** 2)), rel=1e-13)


# ---------------------------------------------------------------- triple bond
def _bond_energy(ff, symbols, d):
    x = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]])
    return orc.evaluate(ff, symbols, x)


def test_triple_bond_closed_form(cho):
    flat = {P_TRIP2: 0.0, P_TRIP3: 0.0, P_TRIP4: 0.0}
    off = _bond_energy(cases.modified(cho, general={**flat, P_TRIP1: 0.0}), ["C", "O"], 1.13)
    assert off["bond_orders"][0, 1] >= 1.0
    for trip1 in (-70.1292, 5.0):
        on = _bond_energy(cases.modified(cho, general={**flat, P_TRIP1: trip1}), ["C", "O"], 1.13)
        # 1 / (1 + 25) * (1 + 1), once per unordered C-O pair
        assert on["components"]["bond"] - off["components"]["bond"] == pytest.approx(trip1 / 13, rel=1e-12)
    # zero below bo = 1
    for d in (1.6, 1.9):
        lo = [_bond_energy(cases.modified(cho, general={**flat, P_TRIP1: t}), ["C", "O"], d) for t in (0.0, 5.0)]
        assert lo[0]["bond_orders"][0, 1] < 1.0
        assert lo[1]["components"]["bond"] == lo[0]["components"]["bond"]
    # only C-O pairs (by mass) unless p(trip) mode 2 allows every pair (general[37])
    cc = [_bond_energy(cases.modified(cho, general={**flat, P_TRIP1: t}), ["C", "C"], 1.2) for t in (0.0, 5.0)]
    assert cc[0]["bond_orders"][0, 1] >= 1.0
    assert cc[1]["components"]["bond"] == cc[0]["components"]["bond"]
    cc2 = [_bond_energy(cases.modified(cho, general={**flat, P_TRIP1: t, 37: 2.0}), ["C", "C"], 1.2)
           for t in (0.0, 5.0)]
    assert cc2[1]["components"]["bond"] - cc2[0]["components"]["bond"] == pytest.approx(5.0 / 13, rel=1e-12)


def test_triple_bond_factors(cho):
    """stab = trip1 e^(-trip2 (b - 2.5)^2) / (1 + 25 e^(trip3 (d_C + d_O)))
    * (e^(-trip4 (T_C - b)) + e^(-trip4 (T_O - b))), b the corrected C-O bond order."""
    g = cho.general
    base = _bond_energy(cases.modified(cho, general={P_TRIP1: 0.0}), ["C", "O"], 1.16)
    full = _bond_energy(cho, ["C", "O"], 1.16)
    b, tot = full["bond_orders"][0, 1], full["total_bond_orders"]
    d = tot - np.array([cho.atoms["C"]["valency"], cho.atoms["O"]["valency"]])
    stab = g[P_TRIP1] * np.exp(-g[P_TRIP2] * (b - 2.5) ** 2) / (1 + 25 * np.exp(g[P_TRIP3] * d.sum()))
    stab *= np.exp(-g[P_TRIP4] * (tot[0] - b)) + np.exp(-g[P_TRIP4] * (tot[1] - b))
    assert full["components"]["bond"] - base["components"]["bond"] == pytest.approx(stab, rel=1e-11)


# ----------------------------------------------------------------- C2 correction
def test_c2_correction_linear_and_thresholded(cho):
    for d in (1.2, 1.25):
        e = [_bond_energy(cases.modified(cho, general={KC2: k}), ["C", "C"], d) for k in (0.0, 35.0, 70.0, 140.0)]
        lp = np.array([r["components"]["lone_pair"] for r in e])
        bo, tot = e[0]["bond_orders"][0, 1], e[0]["total_bond_orders"][0]
        delta = tot - cho.atoms["C"]["valency"]
        vov3 = bo - delta - 0.040 * delta**4
        assert vov3 > 3
        # kc2 (vov3 - 3)^2 on each of the two directed C-C edges
        np.testing.assert_allclose(lp[1:] - lp[0], 2 * np.array([35.0, 70.0, 140.0]) * (vov3 - 3) ** 2, rtol=1e-12)
    for d in (1.3, 1.6):
        e = [_bond_energy(cases.modified(cho, general={KC2: k}), ["C", "C"], d) for k in (0.0, 70.0)]
        bo, tot = e[0]["bond_orders"][0, 1], e[0]["total_bond_orders"][0]
        delta = tot - cho.atoms["C"]["valency"]
        assert bo - delta - 0.040 * delta**4 < 3
        assert e[1]["components"]["lone_pair"] == e[0]["components"]["lone_pair"]
    # across the threshold (vov3 = 3 near d = 1.2714 A) the term goes to zero continuously
    vals, gaps = [], []
    for d in np.linspace(1.266, 1.276, 11):
        e = [_bond_energy(cases.modified(cho, general={KC2: k}), ["C", "C"], d) for k in (0.0, 70.0)]
        vals.append(e[1]["components"]["lone_pair"] - e[0]["components"]["lone_pair"])
        bo, tot = e[0]["bond_orders"][0, 1], e[0]["total_bond_orders"][0]
        delta = tot - cho.atoms["C"]["valency"]
        gaps.append(bo - delta - 0.040 * delta**4 - 3)
    vals, gaps = np.array(vals), np.array(gaps)
    assert (gaps > 0).any() and (gaps < 0).any()
    np.testing.assert_allclose(vals, np.where(gaps > 0, 140.0 * gaps**2, 0.0), rtol=1e-11, atol=0)


# ---------------------------------------------------------------- lammps_angle
HOH = ["H", "O", "H"]


def _angle_gradient(ff, x, lammps):
    model = orc.Model(ff, HOH, orc.build_neighbors(x, ff.general[12]), lammps_forces=lammps)
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    comps, _ = model.components(xt, np.zeros(len(x)))
    (g,) = torch.autograd.grad(comps[orc.COMPONENTS.index("angle")], xt)
    return comps[orc.COMPONENTS.index("angle")].item(), g.numpy()


def test_lammps_angle_matches_fd_above_clamp(ho):
    theta = np.pi - 1e-3
    x = cases.bent_triatomic(theta, (0.97, 0.97))
    _, gl = _angle_gradient(ho, x, True)
    _, gp = _angle_gradient(ho, x, False)
    np.testing.assert_allclose(gl, gp, rtol=1e-9, atol=1e-14)  # -g/sin(arccos c) against -g/sqrt(1 - c^2)
    h = 1e-7
    for atom, axis in ((2, 1), (2, 0), (0, 1)):
        xp, xm = x.copy(), x.copy()
        xp[atom, axis] += h
        xm[atom, axis] -= h
        fd = (_angle_gradient(ho, xp, False)[0] - _angle_gradient(ho, xm, False)[0]) / (2 * h)
        assert fd == pytest.approx(gp[atom, axis], rel=2e-6, abs=1e-8)


def test_lammps_angle_regularised_below_clamp(ho):
    """Below sin(theta) = 1e-5 the theta-part of the gradient is -g / 1e-5 instead of -g / sin:
    it scales by sin(theta) / 1e-5.  The part through the bond orders acts along the bonds,
    so the out-of-line (y) component of atom 2 isolates the theta-part up to |B| sin(theta)."""
    for delta in (1e-6, 3e-6):
        theta = np.pi - delta
        x = cases.bent_triatomic(theta, (0.97, 0.97))
        _, gl = _angle_gradient(ho, x, True)
        _, gp = _angle_gradient(ho, x, False)
        s = np.sin(theta)
        assert gl[2, 1] / gp[2, 1] == pytest.approx(s / 1e-5, rel=1e-3)
        # along the line the bond-order part dominates; the theta-part there is O(|T| sin theta)
        np.testing.assert_allclose(gl[:, 0], gp[:, 0], rtol=0, atol=1e-5)
    # |cos| > 1 - 1e-14: the plain arm clips and its theta-part vanishes; the LAMMPS arm keeps
    # sin(theta) / 1e-5 of the unclipped theta-derivative
    x6 = cases.bent_triatomic(np.pi - 1e-6, (0.97, 0.97))
    _, gp6 = _angle_gradient(ho, x6, False)
    theta = np.pi - 1e-8
    x = cases.bent_triatomic(theta, (0.97, 0.97))
    _, gl = _angle_gradient(ho, x, True)
    _, gp = _angle_gradient(ho, x, False)
    assert abs(gp[2, 1]) < 1e-6 * abs(gp6[2, 1])
    assert gl[2, 1] == pytest.approx(np.sin(theta) / 1e-5 * gp6[2, 1], rel=1e-3)
