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


SHIFT_LIMIT = 511  # image shifts are packed as three 10-bit fields on the device
WRAP_LIMIT = 256  # positions further than this many cells from the cell are wrapped before the device build


def canonical_neighbors(neighbors, atoms, cell=None, pbc=None):
    """Validate a caller's full directed list ``(i, j, S)`` and return it as canonical CSR.

    Same checks, order and messages as the reference's ``Neighbors``
    (src/xreac/neighbors.py:101-131, duplicates and both directions from
    ``_edge_order``, neighbors.py:63-98).  Returns ``row_ptr`` int32 (atoms + 1),
    ``j`` int32 (M,) and ``S`` int64 (M, 3), rows in (i, j, Sx, Sy, Sz) order;
    shifts of any size are accepted, as the reference's overflow-safe order
    does (neighbors.py:63-98).  ``fit_device_range`` then maps them into the
    device's image range.
    """
    boundary = Boundary(cell, pbc)
    if not isinstance(neighbors, (tuple, list)) or len(neighbors) != 3:
        raise ValueError("neighbors must be a tuple (i, j, S)")
    i, j, shifts = (np.asarray(v) for v in neighbors)
    if i.ndim != 1 or j.shape != i.shape or shifts.shape != (len(i), 3):
        raise ValueError("neighbors require i and j with shape (M,) and S with shape (M, 3)")
    if any(v.dtype.kind not in "iu" for v in (i, j, shifts)):
        raise ValueError("Neighbor indices and lattice shifts must be integers")
    if np.any(i < 0) or np.any(j < 0) or np.any(i >= atoms) or np.any(j >= atoms):
        raise ValueError("Neighbor atom indices are out of range")
    if np.any(shifts[:, ~boundary.pbc] != 0):
        raise ValueError("Neighbor shifts must be zero in nonperiodic directions")
    if np.any((i == j) & np.all(shifts == 0, axis=1)):
        raise ValueError("Zero-shift self neighbors are not allowed")
    i = i.astype(np.int64)
    j = j.astype(np.int64)
    s = shifts.astype(np.int64)
    order = np.lexsort((s[:, 2], s[:, 1], s[:, 0], j, i))
    i, j, s = i[order], j[order], s[order]
    forward = np.column_stack((i, j, s))
    if len(i) > 1 and np.any(np.all(forward[1:] == forward[:-1], axis=1)):
        raise ValueError("Duplicate neighbor edges are not allowed")
    # every (i, j, S) needs its (j, i, -S): the reversed set, sorted, must be the same list
    backward = np.column_stack((j, i, -s))
    back = backward[np.lexsort((backward[:, 4], backward[:, 3], backward[:, 2], backward[:, 1], backward[:, 0]))]
    if not np.array_equal(forward, back):
        raise ValueError("neighbors must include both directions: (i, j, S) and (j, i, -S)")
    row_ptr = np.zeros(atoms + 1, dtype=np.int32)
    np.cumsum(np.bincount(i, minlength=atoms), out=row_ptr[1:])
    return row_ptr, j.astype(np.int32), s


def fit_device_range(x, cell, pbc, cutoff, csr=None):
    """Positions (and a supplied canonical CSR) inside the device's image range.

    The device packs image shifts into 10-bit fields (|S_k| <= 511) and builds
    its own lists from positions within 511 cells of the cell.  Moving atom a
    by whole cells, x'_a = x_a - w_a L, leaves every edge vector unchanged when
    its shift becomes S' = S + w_j - w_i, so results do not change (up to the
    rounding of the subtraction).  Positions are wrapped only when some atom
    lies more than WRAP_LIMIT cells out or a supplied shift exceeds 511; after
    wrapping, an edge with |S'_k| > 1 is at least (|S'_k| - 1) h_k long (h_k the
    cell height), and one longer than the cutoff contributes nothing (every
    term is zero there, energy.py:89-103, 105-165, 248-276), so such edges are
    dropped with their reverses.  Returns (x, csr) with csr shifts as int32.
    """
    boundary = Boundary(cell, pbc)
    x = np.asarray(x, dtype=np.float64)
    if csr is None and not boundary.periodic:
        return x, None
    if csr is not None and not boundary.periodic:
        row_ptr, j, s = csr
        return x, (row_ptr, j, np.asarray(s).astype(np.int32))
    flags = boundary.pbc
    frac = x @ boundary.inverse
    far = np.any(np.abs(frac[:, flags]) > WRAP_LIMIT)
    big = csr is not None and len(csr[1]) > 0 and int(np.max(np.abs(csr[2]))) > SHIFT_LIMIT
    if not far and not big:
        if csr is None:
            return x, None
        row_ptr, j, s = csr
        return x, (row_ptr, j, np.asarray(s).astype(np.int32))
    w = np.zeros_like(frac)
    w[:, flags] = np.floor(frac[:, flags])
    w = w.astype(np.int64)
    x = x - w @ boundary.cell
    if csr is None:
        return x, None
    row_ptr, j, s = csr
    rows = np.repeat(np.arange(len(row_ptr) - 1), np.diff(row_ptr))
    s = np.asarray(s, dtype=np.int64) + w[j] - w[rows]
    # lower bound of each edge's length after wrapping; fractional offsets lie in (-1, 1)
    reach = np.max(np.maximum(np.abs(s) - 1 - 1e-9, 0) * boundary.heights, axis=1)
    keep = reach <= cutoff * (1 + 1e-8)
    if np.any(np.abs(s[keep]) > SHIFT_LIMIT):
        raise ValueError(
            f"Neighbor shifts beyond {SHIFT_LIMIT} cells within the cutoff are not supported on the device"
        )
    counts = np.bincount(rows[keep], minlength=len(row_ptr) - 1)
    new_ptr = np.zeros_like(row_ptr)
    np.cumsum(counts, out=new_ptr[1:])
    # within a row and a neighbor j the shift moves by the constant w_j - w_i, so the kept
    # edges stay in canonical order
    return x, (new_ptr, j[keep], s[keep].astype(np.int32))
