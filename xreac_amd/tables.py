"""Flatten a parsed ForceField into the dense device tables of ``include/xreac.h``.

Types are the force field's own atom types in file order.  The reference
instead indexes the labels present in each call (src/xreac/energy.py:73-81);
both select the same parameter values for every atom, pair, angle, torsion and
hydrogen-bond key.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

ATOM_COLUMNS = (
    "r_s", "r_pi", "r_pi_pi", "valency", "valency_e", "valency_boc", "valency_val", "mass", "chi", "eta",
    "gamma", "p_lp2", "p_ovun2", "p_ovun5", "p_val3", "p_val5", "p_hbond",
)
PAIR_COLUMNS = (
    "p_bo1", "p_bo2", "p_bo3", "p_bo4", "p_bo5", "p_bo6", "r_s", "r_p", "r_pp", "ovc", "v13cor",
    "p_boc3", "p_boc4", "p_boc5", "De_s", "De_p", "De_pp", "p_be1", "p_be2", "p_ovun1", "D",
    "alpha", "gamma_w", "r_vdW", "rcore", "ecore", "acore", "gamma",
)
ATOM_WIDTH = len(ATOM_COLUMNS) + 3  # + is_carbon, mass == 12.0, mass == 15.999
PAIR_WIDTH = len(PAIR_COLUMNS)


@dataclass
class Tables:
    labels: tuple
    index: dict
    general: np.ndarray
    atom: np.ndarray
    pair: np.ndarray
    angle_offsets: np.ndarray
    angle_rows: np.ndarray
    torsion: np.ndarray
    torsion_on: np.ndarray
    hbond: np.ndarray
    hbond_on: np.ndarray
    vdw_type: int


def build_tables(ff) -> Tables:
    labels = ff.elements
    nt = len(labels)
    index = {s: k for k, s in enumerate(labels)}
    atom = np.zeros((nt, ATOM_WIDTH))
    for t, s in enumerate(labels):
        a = ff.atoms[s]
        atom[t, : len(ATOM_COLUMNS)] = [float(a[c]) for c in ATOM_COLUMNS]
        atom[t, len(ATOM_COLUMNS)] = 1.0 if s.upper() == "C" else 0.0  # energy.py:233
        atom[t, len(ATOM_COLUMNS) + 1] = 1.0 if a["mass"] == 12.0 else 0.0  # energy.py:239
        atom[t, len(ATOM_COLUMNS) + 2] = 1.0 if a["mass"] == 15.999 else 0.0
    pair = np.zeros((nt * nt, PAIR_WIDTH))
    for ti, s in enumerate(labels):
        for tj, u in enumerate(labels):
            p = ff.pairs[s, u]
            pair[ti * nt + tj] = [float(p[c]) for c in PAIR_COLUMNS]
    # Angles: rows in parse order, kept only when |v1| > 0.001 (energy.py:286-289).
    counts = np.zeros(nt**3, dtype=np.int64)
    rows = []
    for ti, si in enumerate(labels):
        for tj, sj in enumerate(labels):
            for tk, sk in enumerate(labels):
                key = (ti * nt + tj) * nt + tk
                for prm in ff.angles.get((si, sj, sk), []):
                    if abs(prm[1]) > 0.001:
                        rows.append(np.asarray(prm, dtype=float))
                        counts[key] += 1
    angle_offsets = np.concatenate(([0], np.cumsum(counts))).astype(np.int32)
    angle_rows = np.array(rows, dtype=float).reshape(-1, 7)
    # Torsions: active when any of v1, v2, v3, cot1 is nonzero (energy.py:84-87).
    torsion = np.zeros((nt**4, 5))
    torsion_on = np.zeros(nt**4, dtype=np.uint8)
    for (a, b, c, d), prm in ff.torsions.items():
        key = ((index[a] * nt + index[b]) * nt + index[c]) * nt + index[d]
        torsion[key] = prm[:5]
        torsion_on[key] = 1 if np.any(np.asarray(prm)[[0, 1, 2, 4]] != 0) else 0
    # Hydrogen bonds keyed (donor heavy atom, H, acceptor); active when r0 > 0 (energy.py:262-266).
    hbond = np.zeros((nt**3, 4))
    hbond_on = np.zeros(nt**3, dtype=np.uint8)
    for (a, b, c), prm in ff.hydrogen_bonds.items():
        key = (index[a] * nt + index[b]) * nt + index[c]
        hbond[key] = prm
        hbond_on[key] = 1 if prm[0] > 0 else 0
    return Tables(
        labels=labels,
        index=index,
        general=np.ascontiguousarray(ff.general, dtype=float),
        atom=np.ascontiguousarray(atom),
        pair=np.ascontiguousarray(pair),
        angle_offsets=angle_offsets,
        angle_rows=np.ascontiguousarray(angle_rows),
        torsion=np.ascontiguousarray(torsion),
        torsion_on=torsion_on,
        hbond=np.ascontiguousarray(hbond),
        hbond_on=hbond_on,
        vdw_type=int(ff.vdw_type),
    )
