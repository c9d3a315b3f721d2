"""Seeded synthetic systems for the benchmark and parity tests.

water_cluster  - 64 H2O non-periodic cluster (192 atoms, HO.2015)
water_box      - 4096 H2O periodic cubic box at 0.997 g/cm^3 (12,288 atoms,
                 HO.2015) on a jittered 16^3 lattice; ``prerelax`` applies the
                 fixed steepest-descent rule below
cho_mixture    - CH4 / O2 / H2O / CH3OH mixture in a periodic cube (CHO.2008)

Every generator takes a seed and uses numpy.random.default_rng.
"""

from __future__ import annotations

import numpy as np

AVOGADRO_PER_A3 = 0.602214076  # g/cm^3 -> amu / A^3 conversion factor
OH = 0.9572
HOH = np.deg2rad(104.52)


def water_monomer():
    """O at the origin, H atoms in the xy-plane (O-H 0.9572 A, H-O-H 104.52 deg)."""
    half = HOH / 2
    return np.array([[0.0, 0.0, 0.0], [OH * np.cos(half), OH * np.sin(half), 0.0], [OH * np.cos(half), -OH * np.sin(half), 0.0]])


def random_rotation(rng):
    """Uniform SO(3) rotation from a normalised 4-Gaussian quaternion."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array(
        [
            [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
            [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
            [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
        ]
    )


def water_cluster(molecules=64, radius=8.0, min_oo=2.7, jitter=0.02, seed=0):
    """O sites rejection-sampled uniformly in a sphere with a minimum O-O distance."""
    rng = np.random.default_rng(seed)
    sites = []
    while len(sites) < molecules:
        p = rng.uniform(-radius, radius, 3)
        if p @ p > radius * radius:
            continue
        if sites and np.min(np.linalg.norm(np.array(sites) - p, axis=1)) < min_oo:
            continue
        sites.append(p)
    mono = water_monomer()
    x = np.concatenate([mono @ random_rotation(rng).T + s for s in sites])
    x = x + rng.normal(0.0, jitter, x.shape)
    return ["O", "H", "H"] * molecules, x


def water_box_length(molecules=4096, density=0.997, mass=(15.999 + 2 * 1.008)):
    return float((molecules * mass / (density * AVOGADRO_PER_A3)) ** (1.0 / 3.0))


def water_box(side=16, density=0.997, jitter=0.1, seed=1):
    """side^3 waters on a jittered cubic lattice with random orientations, periodic cube."""
    rng = np.random.default_rng(seed)
    n = side**3
    L = water_box_length(n, density)
    a = L / side
    mono = water_monomer()
    mono = mono - mono.mean(axis=0)
    grid = np.array(np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij")).reshape(3, -1).T
    centers = (grid + 0.5) * a + rng.normal(0.0, jitter, (n, 3))
    x = np.concatenate([mono @ random_rotation(rng).T + c for c in centers])
    return ["O", "H", "H"] * n, x, np.diag([L, L, L])


def steepest_descent(evaluate_forces, x, steps=200, gain=0.002, max_step=0.05):
    """Fixed rule: x += clip(gain * F) with each atom's displacement capped at max_step A.

    ``evaluate_forces(x) -> forces`` in kcal/mol/A; gain in A^2 mol/kcal.
    """
    x = np.array(x, dtype=float, copy=True)
    for _ in range(steps):
        f = evaluate_forces(x)
        step = gain * f
        norm = np.linalg.norm(step, axis=1, keepdims=True)
        step = np.where(norm > max_step, step * (max_step / np.maximum(norm, 1e-300)), step)
        x = x + step
    return x


_GAS = {
    # standard gas-phase geometries (A): CH4 (C-H 1.087), O2 (1.208),
    # H2O (0.9572, 104.52 deg), CH3OH (C-O 1.427, O-H 0.956, C-H 1.093)
    "CH4": (["C", "H", "H", "H", "H"], None),
    "O2": (["O", "O"], np.array([[0.0, 0, 0], [1.208, 0, 0]])),
    "H2O": (["O", "H", "H"], None),
    "CH3OH": (["C", "O", "H", "H", "H", "H"], None),
}


def _ch4():
    t = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / np.sqrt(3) * 1.087
    return np.vstack(([0.0, 0, 0], t))


def _methanol():
    c = np.array([0.0, 0, 0])
    o = np.array([1.427, 0, 0])
    h_o = o + 0.956 * np.array([np.cos(np.deg2rad(180 - 108.5)), np.sin(np.deg2rad(180 - 108.5)), 0])
    hs = []
    for phi in (60.0, 180.0, 300.0):
        ang = np.deg2rad(109.5)
        d = np.array([np.cos(ang), np.sin(ang) * np.cos(np.deg2rad(phi)), np.sin(ang) * np.sin(np.deg2rad(phi))])
        hs.append(c + 1.093 * d)
    return np.vstack((c, o, h_o, *hs))


def molecule(name):
    if name == "CH4":
        return ["C", "H", "H", "H", "H"], _ch4()
    if name == "O2":
        return ["O", "O"], _GAS["O2"][1]
    if name == "H2O":
        return ["O", "H", "H"], water_monomer()
    if name == "CH3OH":
        return ["C", "O", "H", "H", "H", "H"], _methanol()
    raise KeyError(name)


def cho_mixture(counts=(("CH4", 300), ("O2", 300), ("H2O", 150), ("CH3OH", 100)), length=40.0, min_heavy=2.5,
                seed=2, max_tries=200000):
    """Randomly placed and oriented molecules in a periodic cube, heavy atoms >= min_heavy apart."""
    rng = np.random.default_rng(seed)
    L = float(length)
    symbols, coords, heavy = [], [], []
    order = [name for name, c in counts for _ in range(c)]
    rng.shuffle(order)
    tries = 0
    for name in order:
        s, m = molecule(name)
        m = m - m.mean(axis=0)
        while True:
            tries += 1
            if tries > max_tries:
                raise RuntimeError("could not place all molecules")
            pos = m @ random_rotation(rng).T + rng.uniform(0, L, 3)
            hv = pos[[k for k, e in enumerate(s) if e != "H"]]
            if heavy:
                H = np.array(heavy)
                d = hv[:, None, :] - H[None, :, :]
                d -= L * np.round(d / L)
                if np.min(np.sqrt(np.sum(d * d, axis=-1))) < min_heavy:
                    continue
            heavy.extend(hv)
            symbols.extend(s)
            coords.append(pos)
            break
    return symbols, np.concatenate(coords), np.diag([L, L, L])
