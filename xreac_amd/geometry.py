"""Host-side cell handling: validation, face heights and the replication count.

Semantics of the reference's ``Boundary`` (src/xreac/geometry.py:15-79):
cell rows are lattice vectors, three lengths mean a diagonal cell, supplying a
cell enables periodicity unless ``pbc`` says otherwise, and the cell must be
finite, nonsingular and right-handed.  The device neighbor build needs no
replicated copies; ``repetitions`` is computed only to report
``cell_repetitions`` and to honour ``max_expanded_atoms`` as the reference does.
"""

from __future__ import annotations

from numbers import Integral

import numpy as np


def validate_expansion_limit(value):
    if isinstance(value, bool) or not isinstance(value, (Integral, np.integer)) or value < 1:
        raise ValueError("max_expanded_atoms must be a positive integer")


class Boundary:
    def __init__(self, cell=None, pbc=None):
        if pbc is None:
            pbc = cell is not None
        flags = np.asarray(pbc)
        if flags.shape not in ((), (3,)) or flags.dtype.kind not in "biu" or not np.isin(flags, [0, 1]).all():
            raise ValueError("pbc must be a boolean or three boolean flags")
        self.pbc = np.broadcast_to(flags.astype(bool), (3,)).copy()
        self.periodic = bool(self.pbc.any())
        self.cell = None
        if not self.periodic:
            return
        if cell is None:
            raise ValueError("A cell is required when pbc is enabled")
        m = np.array(cell, dtype=float, copy=True)
        if m.shape == (3,):
            if np.any(m <= 0):
                raise ValueError("Cell lengths must be positive")
            m = np.diag(m)
        if m.shape != (3, 3) or not np.isfinite(m).all():
            raise ValueError("cell must contain three finite lengths or a finite (3, 3) matrix")
        if np.linalg.det(m) <= 1e-12 * np.prod(np.linalg.norm(m, axis=1)):
            raise ValueError("Periodic cells must be nonsingular and right-handed")
        self.cell = m
        self.inverse = np.linalg.inv(m)
        self.heights = 1.0 / np.linalg.norm(self.inverse, axis=0)

    def repetitions(self, cutoff, atoms, max_expanded_atoms=512):
        """Copies per direction the reference's replicated search would use (geometry.py:58-79)."""
        validate_expansion_limit(max_expanded_atoms)
        reps = np.ones(3, dtype=int)
        if not self.periodic:
            return (1, 1, 1)
        need = np.floor(cutoff * (1 + 1e-12) / self.heights[self.pbc]) + 1
        copies = float(np.prod(need))
        if copies > 1 and copies * atoms > max_expanded_atoms:
            raise ValueError(
                f"Small-cell replication needs {copies * atoms:g} internal atoms, "
                f"exceeding max_expanded_atoms={max_expanded_atoms}; "
                "increase this limit explicitly if memory permits"
            )
        reps[self.pbc] = need.astype(int)
        return tuple(int(v) for v in reps)
