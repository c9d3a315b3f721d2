"""Seeded synthetic systems for the benchmark and parity tests.

water_cluster  - 64 H2O non-periodic cluster (192 atoms, HO.2015)
water_box      - 4096 H2O periodic cubic box at 0.997 g/cm^3 (12,288 atoms,
                 HO.2015) on a jittered 16^3 lattice; ``prerelax`` applies the
                 fixed steepest-descent rule below
cho_mixture    - CH4 / O2 / H2O / CH3OH mixture in a periodic cube (CHO.2008)

Every generator takes a seed and uses numpy.random.default_rng.
"""

from __future__ import annotations

import os

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


def water_cluster(molecules=64, radius=8.0, min_oo=2.7, jitter=0.02, seed=0, max_tries=10**7):
    """O sites rejection-sampled uniformly in a sphere with a minimum O-O distance."""
    rng = np.random.default_rng(seed)
    sites = np.empty((molecules, 3))
    count = 0
    tries = 0
    while count < molecules:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place all water molecules; enlarge the radius")
        p = rng.uniform(-radius, radius, 3)
        if p @ p > radius * radius:
            continue
        if count and np.min(np.linalg.norm(sites[:count] - p, axis=1)) < min_oo:
            continue
        sites[count] = p
        count += 1
    mono = water_monomer()
    x = np.concatenate([mono @ random_rotation(rng).T + s for s in sites])
    x = x + rng.normal(0.0, jitter, x.shape)
    return ["O", "H", "H"] * molecules, x


KB_KCAL = 0.0019872043  # kcal/mol/K
ACCEL = 4.184e-4  # (kcal/mol/A/amu) -> A/fs^2


def maxwell_boltzmann(masses, temperature=300.0, seed=0):
    """Velocities (A/fs) drawn at ``temperature`` K with zero total momentum."""
    rng = np.random.default_rng(seed)
    m = np.asarray(masses, dtype=np.float64)
    v = rng.normal(size=(len(m), 3)) * np.sqrt(KB_KCAL * temperature * ACCEL / m)[:, None]
    if len(m) > 1:
        v -= (m[:, None] * v).sum(axis=0) / m.sum()
    return v


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


PRERELAXED_BOX = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data",
                              "water_box_12288_prerelaxed.npz")


def prerelaxed_water_box(path=PRERELAXED_BOX):
    """The 12,288-atom box after its 200-step steepest-descent pre-relax (data/make_water_box_prerelaxed.py):
    symbols, positions (12,288 x 3) and the cubic cell."""
    data = np.load(path)
    x = np.array(data["positions"], dtype=np.float64)
    return ["O", "H", "H"] * (len(x) // 3), x, np.array(data["cell"], dtype=np.float64)


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


# ------------------------------------------------------ serving stream
# Mixed-size request stream: atom counts log-uniform in [3, 1536]; 60 % water
# clusters (HO.2015), 30 % CHO mixtures (CHO.2008), 10 % water clusters with
# one H3O+ or OH- (net charge +-1); 20 % of all requests periodic cubic cells.
# Densities: water clusters as the 64-molecule, 8 A cluster (33.5 A^3 per
# molecule); periodic water 0.997 g/cm^3; CHO at the number density of the
# 3150-atom, 40 A mixture (0.0492 atoms/A^3), clusters and cells alike.
CHO_DENSITY = 3150 / 40.0**3
CHO_RATIO = (("CH4", 300), ("O2", 300), ("H2O", 150), ("CH3OH", 100))
CHO_SIZES = {"CH4": 5, "O2": 2, "H2O": 3, "CH3OH": 6}


def water_cluster_of(molecules, seed, min_oo=2.7):
    """Cluster at the density of the 64-molecule, 8 A cluster (radius at least 1.5 min O-O, so that a
    few-molecule cluster always has room for its sites)."""
    radius = max(8.0 * (molecules / 64.0) ** (1.0 / 3.0), 1.5 * min_oo)
    return water_cluster(molecules=molecules, radius=radius, min_oo=min_oo, seed=seed)


def water_cell(molecules, density=0.997, jitter=0.1, seed=0):
    """Periodic cubic cell of water at ``density`` (molecules on a jittered lattice, random orientations)."""
    rng = np.random.default_rng(seed)
    L = water_box_length(molecules, density)
    side = int(np.ceil(molecules ** (1.0 / 3.0) - 1e-9))
    sites = np.array(np.meshgrid(*[np.arange(side)] * 3, indexing="ij")).reshape(3, -1).T
    sites = sites[rng.permutation(len(sites))[:molecules]]
    mono = water_monomer()
    x = np.concatenate([mono @ random_rotation(rng).T + (s + 0.5) * (L / side) for s in sites])
    x = x + rng.normal(0.0, jitter, x.shape)
    return ["O", "H", "H"] * molecules, x, np.diag([L, L, L])


def ion_cluster(molecules, sign, seed):
    """Water cluster in which one molecule is H3O+ (sign +1) or OH- (sign -1)."""
    rng = np.random.default_rng(seed + 7919)
    s, x = water_cluster_of(molecules, seed)
    s = list(s)
    k = int(rng.integers(molecules))
    o, h1, h2 = x[3 * k], x[3 * k + 1], x[3 * k + 2]
    if sign > 0:
        bis = (h1 - o) + (h2 - o)
        away = -bis / np.linalg.norm(bis)
        perp = np.cross(h1 - o, h2 - o)
        perp /= np.linalg.norm(perp)
        d = 0.55 * away + 0.83 * perp
        d /= np.linalg.norm(d)
        return s + ["H"], np.vstack((x, o + 0.98 * d)), 1.0
    keep = [i for i in range(len(s)) if i != 3 * k + 2]
    return [s[i] for i in keep], x[keep], -1.0


def cho_counts(atoms, rng):
    """Molecule counts in the CH4:O2:H2O:CH3OH = 300:300:150:100 ratio totalling about ``atoms`` atoms."""
    names = [n for n, _ in CHO_RATIO]
    weights = np.array([c for _, c in CHO_RATIO], dtype=float)
    weights /= weights.sum()
    counts = dict.fromkeys(names, 0)
    total = 0
    while total < atoms:
        name = names[int(rng.choice(len(names), p=weights))]
        if total + CHO_SIZES[name] > atoms and total > 0:
            break
        counts[name] += 1
        total += CHO_SIZES[name]
    return tuple((n, c) for n, c in counts.items() if c)


def cho_cluster(counts, radius, min_heavy=2.5, seed=0, max_tries=200000):
    """Randomly placed and oriented molecules inside a sphere, heavy atoms >= min_heavy apart."""
    rng = np.random.default_rng(seed)
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
            c = rng.uniform(-radius, radius, 3)
            if c @ c > radius * radius:
                continue
            pos = m @ random_rotation(rng).T + c
            hv = pos[[k for k, e in enumerate(s) if e != "H"]]
            if heavy and np.min(np.linalg.norm(hv[:, None, :] - np.array(heavy)[None], axis=-1)) < min_heavy:
                continue
            heavy.extend(hv)
            symbols.extend(s)
            coords.append(pos)
            break
    return symbols, np.concatenate(coords)


def stream_plan(count=20000, seed=4, min_atoms=3, max_atoms=1536, md_fraction=0.2, md_steps=50, dt=0.25,
                temperature=300.0):
    """Cheap, sequential draws of the mixed stream: one spec dict per request (see ``build_request``)."""
    rng = np.random.default_rng(seed)
    plan = []
    for k in range(count):
        atoms = int(np.round(np.exp(rng.uniform(np.log(min_atoms), np.log(max_atoms)))))
        family = rng.uniform()
        periodic = bool(rng.uniform() < 0.2)
        md = bool(rng.uniform() < md_fraction)
        sub = int(rng.integers(2**31))
        sign = (1 if rng.uniform() < 0.5 else -1) if family >= 0.9 else 0
        plan.append(dict(rid=k, atoms=atoms, family=float(family), periodic=periodic, md_steps=md_steps if md else 0,
                         dt=dt, temperature=temperature, seed=sub, sign=sign))
    return plan


def build_request(spec):
    """Geometry of one planned request: dict(rid, force_field, symbols, positions, cell, pbc, total_charge,
    md_steps, dt, temperature, velocity_seed)."""
    atoms, family, periodic, seed = spec["atoms"], spec["family"], spec["periodic"], spec["seed"]
    q, cell = 0.0, None
    if family < 0.3:
        ff = "ffield.reax.CHO.2008"
        counts = cho_counts(atoms, np.random.default_rng(seed))
        n = sum(CHO_SIZES[name] * c for name, c in counts)
        if periodic:
            L = max((n / CHO_DENSITY) ** (1.0 / 3.0), 8.0)
            s, x, cell = cho_mixture(counts=counts, length=L, seed=seed)
        else:
            s, x = cho_cluster(counts, radius=max((3 * n / (4 * np.pi * CHO_DENSITY)) ** (1.0 / 3.0), 4.0), seed=seed)
    else:
        ff = "ffield.reax.HO.2015"
        molecules = max(1, int(round(atoms / 3)))
        if family < 0.9:
            if periodic:
                s, x, cell = water_cell(molecules, seed=seed)
            else:
                s, x = water_cluster_of(molecules, seed)
        else:
            s, x, q = ion_cluster(max(2, molecules), spec["sign"], seed)
    return dict(rid=spec["rid"], force_field=ff, symbols=list(s), positions=np.asarray(x, dtype=np.float64),
                cell=cell, pbc=(True, True, True) if cell is not None else (False, False, False), total_charge=q,
                md_steps=spec["md_steps"], dt=spec["dt"], temperature=spec["temperature"], velocity_seed=seed)


def request_stream(count=20000, seed=4, workers=1, **kw):
    """The seeded mixed request stream; ``workers`` > 1 builds geometries in a process pool."""
    plan = stream_plan(count, seed, **kw)
    if workers <= 1:
        return [build_request(p) for p in plan]
    from concurrent.futures import ProcessPoolExecutor

    with ProcessPoolExecutor(max_workers=workers) as pool:
        return list(pool.map(build_request, plan, chunksize=16))


# ------------------------------------------------------ anatase + water slab
ANATASE_A, ANATASE_C, ANATASE_U = 3.785, 9.514, 0.208


def anatase_cell(a=ANATASE_A, c=ANATASE_C, u=ANATASE_U):
    """Conventional anatase cell (I4_1/amd, origin choice 1): 4 Ti at 4a, 8 O at 8e (0, 0, u)."""
    ti = [(0, 0, 0), (0.5, 0.5, 0.5), (0, 0.5, 0.25), (0.5, 0, 0.75)]
    o = [(0, 0, u), (0, 0, -u), (0.5, 0.5, 0.5 + u), (0.5, 0.5, 0.5 - u),
         (0, 0.5, 0.25 + u), (0, 0.5, 0.25 - u), (0.5, 0, 0.75 - u), (0.5, 0, 0.75 + u)]
    frac = np.array(ti + o, dtype=np.float64) % 1.0
    return ["Ti"] * 4 + ["O"] * 8, frac * np.array([a, a, c])


def anatase_water_slab(nx=16, ny=16, nz=5, waters=3000, water_thickness=25.0, vacuum=15.0, gap=1.8, jitter=0.02,
                       seed=3):
    """Anatase (001) slab of nx*ny*nz conventional cells under a water layer, then vacuum, 3-D periodic.

    Defaults: 15,360 slab atoms + 3000 H2O (about 1.0 g/cm^3 over 25 A) + 15 A vacuum = 24,360 atoms.
    Slab atoms get N(0, jitter) displacements: the perfect lattice has exactly linear O-Ti-O, which
    the torsion term rejects as collinear (energy.py:320-370).
    """
    rng = np.random.default_rng(seed)
    s0, x0 = anatase_cell()
    shifts = np.array(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")).reshape(3, -1).T
    x = (x0[None, :, :] + (shifts * np.array([ANATASE_A, ANATASE_A, ANATASE_C]))[:, None, :]).reshape(-1, 3)
    x = x + rng.normal(0.0, jitter, x.shape)
    symbols = s0 * len(shifts)
    lx, ly, top = nx * ANATASE_A, ny * ANATASE_A, x[:, 2].max()
    # water on a jittered lattice filling [top + gap, top + water_thickness]
    height = water_thickness - gap
    per = (lx * ly * height / waters) ** (1.0 / 3.0)
    mx, my, mz = max(1, int(lx / per)), max(1, int(ly / per)), max(1, int(np.ceil(height / per)))
    while mx * my * mz < waters:
        mz += 1
    sites = np.array(np.meshgrid(np.arange(mx), np.arange(my), np.arange(mz), indexing="ij")).reshape(3, -1).T
    sites = sites[rng.permutation(len(sites))[:waters]]
    step = np.array([lx / mx, ly / my, (height - 1.0) / max(mz - 1, 1)])
    centers = sites * step + np.array([0.5 * step[0], 0.5 * step[1], top + gap + 0.5])
    mono = water_monomer()
    wx = np.concatenate([mono @ random_rotation(rng).T + c for c in centers])
    wx = wx + rng.normal(0.0, 0.05, wx.shape)
    cell = np.diag([lx, ly, top + water_thickness + vacuum])
    return symbols + ["O", "H", "H"] * waters, np.vstack((x, wx)), cell
