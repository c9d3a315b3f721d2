"""Deterministic geometries used across the tests.

Geometries restate the reference's fixture builders: ``tests/cases.py:6-19``,
``examples/water_cluster.py:22-36`` (``water_cases``),
``examples/periodic_water.py:27-61`` and ``examples/small_cells.py:24-45``.
"""

from itertools import product

import numpy as np


def zno_cluster(n, spacing=2.15):
    side = int(np.ceil(n ** (1 / 3)))
    ijk = np.array(list(np.ndindex(side, side, side)))[:n]
    rng = np.random.default_rng(1729 + n)
    symbols = ["Zn" if row.sum() % 2 == 0 else "O" for row in ijk]
    return symbols, ijk * spacing + rng.normal(0, 0.025, (n, 3))


ZNO = {
    "zn_atom": (["Zn"], np.array([[0.0, 0.0, 0.0]])),
    "zno": (["Zn", "O"], np.array([[0, 0, 0], [1.9, 0, 0]], dtype=float)),
    "o4": (["O"] * 4, np.array([[-1.1, 0.6, 0.2], [0, 0, 0], [1.3, 0.1, 0], [1.8, 1.2, 0.5]])),
    "cluster20": zno_cluster(20),
}


def water_cases():
    monomer = np.array([[0.0, 0.0, 0.0], [0.97, 0.0, 0.0], [-0.243, 0.94, 0.0]])
    half = np.deg2rad(104.5 / 2)
    acceptor = np.array(
        [[0.0, 0.0, 0.0], [0.97 * np.cos(half), 0.97 * np.sin(half), 0.0], [0.97 * np.cos(half), -0.97 * np.sin(half), 0.0]]
    ) + [2.85, 0.15, 0.12]
    distorted = np.vstack((monomer, acceptor))
    distorted[1] += [0.09, 0.04, -0.02]
    distorted[4] += [-0.04, 0.03, 0.06]
    return {"monomer": (["O", "H", "H"], monomer), "distorted_dimer": (["O", "H", "H"] * 2, distorted)}


def water_box(counts=(4, 4, 4), cell=None, pbc=(True, True, True)):
    cell = np.diag([12.48] * 3) if cell is None else np.asarray(cell, dtype=float)
    rng = np.random.default_rng(731)
    monomer = water_cases()["monomer"][1]
    positions = []
    for index in product(*(range(n) for n in counts)):
        center = ((np.asarray(index) + 0.12) / counts) @ cell + rng.normal(0, 0.04, 3)
        rotation, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        positions.extend(monomer @ rotation + center)
    positions = np.asarray(positions)
    fractional = positions @ np.linalg.inv(cell)
    fractional[:, list(pbc)] %= 1
    return ["O", "H", "H"] * int(np.prod(counts)), fractional @ cell


def periodic_cases(include_bulk=False):
    cell = np.diag([12.0] * 3)
    symbols, dimer = water_cases()["distorted_dimer"]
    boundary = (dimer + [11.5, 11.6, 11.8]) % 12
    tilted = np.array([[12.0, 0, 0], [3.0, 12.0, 0], [1.0, 2.0, 12.0]])
    rotation, _ = np.linalg.qr(np.random.default_rng(81).normal(size=(3, 3)))
    tilted = tilted @ rotation
    sparse_symbols, slab = water_box((2, 2, 2), tilted, pbc=(True, True, False))
    cases = {
        "boundary_dimer": (symbols, boundary, cell, [True] * 3),
        "partial_pbc_water": (sparse_symbols, slab, tilted, [True, True, False]),
    }
    if include_bulk:
        s, x = water_box()
        cases["bulk_water_192"] = (s, x, np.diag([12.48] * 3), [True] * 3)
    return cases


def small_cell_cases():
    water = "ffield.reax.HO.2015"
    symbols, x = water_cases()["monomer"]
    cases = {"water_4A": (water, symbols, x, np.diag([4.0] * 3), [True] * 3)}
    rotation, _ = np.linalg.qr(np.random.default_rng(81).normal(size=(3, 3)))
    tilted = np.array([[4.5, 0, 0], [1.0, 5.0, 0], [0.6, 0.4, 5.5]]) @ rotation
    cases["partial_pbc_water"] = (water, symbols, x @ rotation, tilted, [True, True, False])
    cases["zno_4A"] = ("ffield.reax.ZnOH.2010", ["Zn", "O"], np.array([[0.0, 0, 0], [1.9, 0.1, 0.2]]), np.diag([4.0] * 3), [True] * 3)
    cases["zinc_chain"] = ("ffield.reax.ZnOH.2010", ["Zn"], np.zeros((1, 3)), np.diag([2.5, 12.0, 12.0]), [True, False, False])
    return cases


def carbon_cases():
    return {
        "methane": (["C"] + ["H"] * 4, np.vstack(([0, 0, 0], 0.63 * np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]])))),
        "carbon_monoxide": (["C", "O"], np.array([[0, 0, 0], [1.15, 0, 0]], dtype=float)),
        "carbon_dimer": (["C", "C"], np.array([[0, 0, 0], [1.2, 0, 0]], dtype=float)),
    }


def periodic_carbon_chain():
    return (
        "ffield.reax.CHO.2008",
        ["C"] * 3,
        np.array([[0.0, 0, 0], [1.5, 0.4, 0.2], [3.0, -0.2, 0.8]]),
        np.diag([4.5, 12.0, 12.0]),
        [True, False, False],
    )


def charged_cases():
    symbols, water = water_cases()["monomer"]
    cases = {
        "hydroxide": (symbols[:2], water[:2].copy(), None, False, -1.0),
        "hydronium": (symbols + ["H"], np.vstack((water, [-0.3, -0.45, 0.85])), None, False, 1.0),
    }
    for name in ("water_4A", "partial_pbc_water"):
        _, s, x, cell, pbc = small_cell_cases()[name]
        cases["charged_" + name] = (s, x, cell, pbc, -1.0)
    return cases


# ------------------------------------------------ closed-form pin geometries
def modified(ff, general=None, atoms=None, angles=None, torsions=None):
    """Copy of a ForceField with edited parameters.

    general: {index: value}; atoms: {label: {key: value}}; angles: {(a, b, c): [row, ...]};
    torsions: {(a, b, c, d): row}.  Rows follow the file's column order.
    """
    from copy import deepcopy
    from dataclasses import replace

    g = ff.general.copy()
    for k, v in (general or {}).items():
        g[k] = v
    at = deepcopy(ff.atoms)
    for label, kv in (atoms or {}).items():
        at[label].update(kv)
    an = dict(ff.angles)
    for key, rows in (angles or {}).items():
        an[key] = [np.asarray(r, dtype=float) for r in rows]
    to = dict(ff.torsions)
    for key, row in (torsions or {}).items():
        to[key] = np.asarray(row, dtype=float)
    return replace(ff, general=g, atoms=at, angles=an, torsions=to)


def dihedral_chain(omega, lengths=(1.09, 1.54, 1.09), angles=(125.0, 135.0)):
    """Four atoms i-j-k-l with bond lengths |ij|, |jk|, |kl|, bond angles ijk and jkl
    (degrees) and IUPAC dihedral omega (radians; 0 = cis)."""
    t1, t2 = np.radians(angles)
    j = np.zeros(3)
    k = np.array([0.0, 0.0, lengths[1]])
    i = j + lengths[0] * np.array([np.sin(t1), 0.0, np.cos(t1)])
    l = k + lengths[2] * np.array([np.sin(t2) * np.cos(omega), np.sin(t2) * np.sin(omega), -np.cos(t2)])
    return np.array([i, j, k, l])


def iupac_dihedral(p):
    """Signed dihedral of four points (atan2 form, 0 = cis)."""
    b0, b1, b2 = p[0] - p[1], p[2] - p[1], p[3] - p[2]
    b1 = b1 / np.linalg.norm(b1)
    v = b0 - np.dot(b0, b1) * b1
    w = b2 - np.dot(b2, b1) * b1
    return np.arctan2(np.dot(np.cross(b1, v), w), np.dot(v, w))


def bent_triatomic(theta, lengths=(1.4, 1.4)):
    """End-centre-end with the given bond lengths and angle (radians) at the centre (atom 1)."""
    return np.array([[lengths[0], 0.0, 0.0], [0.0, 0.0, 0.0],
                     [lengths[1] * np.cos(theta), lengths[1] * np.sin(theta), 0.0]])
