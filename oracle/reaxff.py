"""CPU FP64 oracle for the ReaxFF energy, QEq and fixed-charge force path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: ``tests/``,
``__graft_entry__.smoke`` and ``bench.py``'s ``cpu_baseline`` leg import it.
The product package ``xreac_amd`` never imports, links or calls it.

It restates ``src/xreac/energy.py`` and ``src/xreac/calculator.py`` of
hsulab/xreac 0.10.0 term by term (line numbers below refer to those files),
with torch-CPU float64 in the role HIPS Autograd plays there.  The three
primitives whose derivative rules matter are rebuilt as custom Functions:

* ``scatter_sum``  - ``np.bincount`` forward, gather backward (neighbors.py:56-62);
* ``lammps_angle`` - ``arccos(clip(c))`` with derivative ``-g/max(sin, 1e-5)``
  (energy.py:49-55);
* ``autograd_clip`` - HIPS Autograd's clip rule: zero gradient when the output
  sits on a bound (used by energy.py:273 and 304).

Enumeration of angle / torsion / hydrogen-bond items is vectorised with NumPy
but emits items in the reference's loop order (energy.py:251-266, 281-289,
323-344).  QEq is the reference's dense KKT LU solve (energy.py:89-103).
"""

from __future__ import annotations

import numpy as np
import torch

from .geometry import boundary, canonical_edges, image_neighbors, shift_offsets

C_ELE = 332.06371
QEQ_COULOMB = 14.4
SELF_CONVERSION = 23.02
THB_CUT = 0.001
THB_PRODUCT_CUT = 0.00001
BOND_CUT = 5.0
HBOND_CUT = 7.5
HBOND_THRESHOLD = 0.01
BOND_GRAPH_CUT = 0.3
COMPONENTS = (
    "bond",
    "atom",
    "lone_pair",
    "molecule",
    "angle",
    "penalty",
    "angle_conjugation",
    "hydrogen_bond",
    "torsion",
    "conjugation",
    "vdw",
    "coulomb",
    "electric_field",
    "qeq",
)
F64 = torch.float64


# --------------------------------------------------------------------- primitives
class _ScatterSum(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, indices, size):
        ctx.save_for_backward(indices)
        out = np.bincount(indices.numpy(), weights=values.detach().numpy(), minlength=size).astype(float)
        return torch.from_numpy(out)

    @staticmethod
    def backward(ctx, g):
        (indices,) = ctx.saved_tensors
        return g[indices], None, None


def scatter_sum(values, indices, size):
    """Sum edge values into bins; repeated indices accumulate (neighbors.py:56-62)."""
    if not torch.is_tensor(indices):
        indices = torch.as_tensor(np.asarray(indices, dtype=np.int64))
    return _ScatterSum.apply(values, indices, int(size))


class _LammpsAngle(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cosine):
        angle = torch.arccos(torch.clamp(cosine, -1.0, 1.0))
        ctx.save_for_backward(angle)
        return angle

    @staticmethod
    def backward(ctx, g):
        (angle,) = ctx.saved_tensors
        return -g / torch.clamp(torch.sin(angle), min=1e-5)


def lammps_angle(cosine):
    return _LammpsAngle.apply(cosine)


class _AutogradClip(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, lo, hi):
        ans = torch.clamp(x, lo, hi)
        ctx.save_for_backward(ans)
        ctx.lo, ctx.hi = lo, hi
        return ans

    @staticmethod
    def backward(ctx, g):
        (ans,) = ctx.saved_tensors
        return g * ((ans != ctx.lo) & (ans != ctx.hi)).to(g.dtype), None, None


def autograd_clip(x, lo, hi):
    return _AutogradClip.apply(x, lo, hi)


class _DenseSolve(torch.autograd.Function):
    """LAPACK gesv through NumPy, as ``np.linalg.solve`` (energy.py:102), differentiable."""

    @staticmethod
    def forward(ctx, a, b):
        x = torch.from_numpy(np.linalg.solve(a.detach().numpy(), b.detach().numpy()))
        ctx.save_for_backward(a, x)
        return x

    @staticmethod
    def backward(ctx, g):
        a, x = ctx.saved_tensors
        gb = torch.from_numpy(np.linalg.solve(a.detach().numpy().T, g.numpy()))
        return -torch.outer(gb, x), gb


def positive_power(x, p):
    """x**p for x > 0 else 0, with a finite derivative everywhere (energy.py:44-46)."""
    safe = torch.where(x > 0, x, torch.ones_like(x))
    return torch.where(x > 0, safe**p, torch.zeros_like(x))


def _t(a):
    return torch.as_tensor(np.asarray(a, dtype=np.float64))


def _where(cond, a, b):
    cond = torch.as_tensor(cond)
    if not torch.is_tensor(a):
        a = torch.tensor(float(a), dtype=F64)
    if not torch.is_tensor(b):
        b = torch.tensor(float(b), dtype=F64)
    return torch.where(cond, a, b)


# ------------------------------------------------------------------- edge tables
class Edges:
    """Validated canonical directed list (neighbors.py:101-141)."""

    def __init__(self, neighbors, atoms, cell=None, pbc=None):
        lattice, flags, periodic = boundary(cell, pbc)
        self.cell = lattice if periodic else np.eye(3)
        self.n = atoms
        if not isinstance(neighbors, (tuple, list)) or len(neighbors) != 3:
            raise ValueError("neighbors must be a tuple (i, j, S)")
        i, j, shifts = (np.asarray(v) for v in neighbors)
        if i.ndim != 1 or j.shape != i.shape or shifts.shape != (len(i), 3):
            raise ValueError("neighbors require i and j with shape (M,) and S with shape (M, 3)")
        if any(v.dtype.kind not in "iu" for v in (i, j, shifts)):
            raise ValueError("Neighbor indices and lattice shifts must be integers")
        if np.any(i < 0) or np.any(j < 0) or np.any(i >= atoms) or np.any(j >= atoms):
            raise ValueError("Neighbor atom indices are out of range")
        if np.any(shifts[:, ~flags] != 0):
            raise ValueError("Neighbor shifts must be zero in nonperiodic directions")
        if np.any((i == j) & np.all(shifts == 0, axis=1)):
            raise ValueError("Zero-shift self neighbors are not allowed")
        order, self.reverse = canonical_edges(i, j, shifts, atoms)
        self.i = i[order].astype(np.int64)
        self.j = j[order].astype(np.int64)
        self.shifts = shifts[order].astype(np.int64)
        self.offsets = shift_offsets(self.shifts, self.cell) if periodic else np.zeros((len(self.i), 3))
        self.half = np.arange(len(self.i)) < self.reverse
        counts = np.bincount(self.i, minlength=self.n)
        self.row_ptr = np.concatenate(([0], np.cumsum(counts)))
        self.ti = torch.as_tensor(self.i)
        self.tj = torch.as_tensor(self.j)

    def geometry(self, x):
        v = x[self.tj] - x[self.ti] + _t(self.offsets)
        r = torch.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
        return v, r

    def atom_sum(self, values):
        return scatter_sum(values, self.ti, self.n)

    def pair_sum(self, values):
        return scatter_sum(values, self.ti * self.n + self.tj, self.n * self.n).reshape(self.n, self.n)


def _rows_of(indices_sorted_by_row, row_of_edge, n):
    """Split an ascending edge-index array into per-atom rows."""
    counts = np.bincount(row_of_edge, minlength=n)
    return np.split(indices_sorted_by_row, np.cumsum(counts)[:-1])


# ------------------------------------------------------------------------- model
class Model:
    """Restatement of ``EnergyModel`` (energy.py:58-370)."""

    def __init__(self, ff, symbols, neighbors, cell=None, pbc=None, *, total_charge=0, lammps_forces=True):
        ff.validate_model(symbols)
        self.ff, self.symbols, self.n = ff, tuple(symbols), len(symbols)
        self.total_charge = total_charge
        self.lammps_forces = lammps_forces
        self.g, self.vdw_type = ff.general, ff.vdw_type
        self.edges = Edges(neighbors, self.n, cell, pbc)
        self.i, self.j = self.edges.i, self.edges.j
        labels = tuple(dict.fromkeys(symbols))
        index = {s: k for k, s in enumerate(labels)}
        self.atom_types = np.array([index[s] for s in symbols], dtype=np.int64)
        self.labels = labels
        pair_types = self.atom_types[self.i] * len(labels) + self.atom_types[self.j]
        self.a = {k: np.array([ff.atoms[s][k] for s in labels])[self.atom_types] for k in ff.atoms[symbols[0]]}
        self.ta = {k: _t(v) for k, v in self.a.items() if k != "p_hbond"}
        pkeys = ff.pairs[symbols[0], symbols[0]].keys()
        self._pair_table = {k: np.array([ff.pairs[s, t][k] for s in labels for t in labels]) for k in pkeys}
        self._pair_types = pair_types
        self._p = {}
        self.torsion_types = {}
        for (a, b, c, d), prm in ff.torsions.items():
            if all(s in index for s in (a, b, c, d)) and np.any(prm[[0, 1, 2, 4]] != 0):
                self.torsion_types.setdefault((b, c), {}).setdefault(a, {})[d] = prm

    def p(self, key, sel=None):
        """Per-edge parameter gathered lazily (energy.py:78-81)."""
        if key not in self._p:
            self._p[key] = self._pair_table[key][self._pair_types]
        v = self._p[key]
        return v if sel is None else v[sel]

    # ------------------------------------------------------------ QEq (89-103)
    def electrostatics(self, r, fixed_charges=None):
        rc = self.g[12]
        u = torch.minimum(r / rc, torch.tensor(1.0, dtype=F64))
        taper = torch.where(r < rc, 1 + u**4 * (-35 + u * (84 + u * (-70 + 20 * u))), torch.zeros_like(r))
        shield = taper / (r**3 + _t(self.p("gamma"))) ** (1 / 3)
        if fixed_charges is not None:
            return torch.as_tensor(np.asarray(fixed_charges, dtype=float)), taper, shield
        h = QEQ_COULOMB * self.edges.pair_sum(shield) + torch.diag(self.ta["eta"])
        n = self.n
        kkt = torch.cat(
            (
                torch.cat((h, torch.ones((n, 1), dtype=F64)), dim=1),
                torch.cat((torch.ones((1, n), dtype=F64), torch.zeros((1, 1), dtype=F64)), dim=1),
            ),
            dim=0,
        )
        rhs = torch.cat((-self.ta["chi"], torch.tensor([float(self.total_charge)], dtype=F64)))
        q = _DenseSolve.apply(kkt, rhs)[:n]
        return q, taper, shield

    # ---------------------------------------------------- bond orders (105-165)
    def bond_orders(self, r):
        a, g = self.a, self.g
        sel = np.flatnonzero(r.detach().numpy() <= min(BOND_CUT, g[12]))
        rs, i, j = r[sel], self.i[sel], self.j[sel]
        P = {k: self.p(k, sel) for k in (
            "p_bo1", "p_bo2", "p_bo3", "p_bo4", "p_bo5", "p_bo6", "r_s", "r_p", "r_pp",
            "ovc", "v13cor", "p_boc3", "p_boc4", "p_boc5")}
        cutoff = 0.01 * g[29]
        sigma_ok = (a["r_s"][i] > 0) & (a["r_s"][j] > 0)
        pi_ok = (a["r_pi"][i] > 0) & (a["r_pi"][j] > 0)
        pp_ok = (a["r_pi_pi"][i] > 0) & (a["r_pi_pi"][j] > 0)
        sigma = _where(sigma_ok, (1 + cutoff) * torch.exp(_t(P["p_bo1"]) * (rs / _t(np.where(sigma_ok, P["r_s"], 1))) ** _t(P["p_bo2"])), 0)
        pi = _where(pi_ok, torch.exp(_t(P["p_bo3"]) * (rs / _t(np.where(pi_ok, P["r_p"], 1))) ** _t(P["p_bo4"])), 0)
        pp = _where(pp_ok, torch.exp(_t(P["p_bo5"]) * (rs / _t(np.where(pp_ok, P["r_pp"], 1))) ** _t(P["p_bo6"])), 0)
        raw = sigma + pi + pp
        mask = raw.detach().numpy() >= cutoff
        corrected = mask & ((P["ovc"] >= 0.001) | (P["v13cor"] >= 0.001))
        for key in ("p_boc3", "p_boc4", "p_boc5"):
            if np.any(corrected & ~np.isfinite(P[key])):
                raise ValueError(f"Invalid {key} mixing for an active bond-order correction")
            P[key] = np.where(corrected, P[key], 0.0)
        bo = _where(mask, raw - cutoff, 0)
        pi, pp = _where(mask, pi, 0), _where(mask, pp, 0)
        ti = torch.as_tensor(i)
        tj = torch.as_tensor(j)
        total = scatter_sum(bo, ti, self.n)
        d = total - self.ta["valency"]
        db = total - self.ta["valency_val"]
        f2 = torch.exp(-g[0] * d[ti]) + torch.exp(-g[0] * d[tj])
        f3 = -torch.log(0.5 * (torch.exp(-g[1] * d[ti]) + torch.exp(-g[1] * d[tj]))) / g[1]
        vi, vj = _t(a["valency"][i]), _t(a["valency"][j])
        f1 = _where(P["ovc"] >= 0.001, 0.5 * ((vi + f2) / (vi + f2 + f3) + (vj + f2) / (vj + f2 + f3)), 1)
        b3, b4, b5 = _t(P["p_boc3"]), _t(P["p_boc4"]), _t(P["p_boc5"])
        f4 = 1 / (1 + torch.exp(-(b4 * bo**2 - db[ti]) * b3 + b5))
        f5 = 1 / (1 + torch.exp(-(b4 * bo**2 - db[tj]) * b3 + b5))
        corr = f1 * _where(P["v13cor"] >= 0.001, f4 * f5, 1)
        bo, pi, pp = bo * corr, pi * corr * f1, pp * corr * f1
        tsel = torch.as_tensor(sel)
        m = len(self.i)
        outs = []
        for v in (bo, bo - pi - pp, pi, pp):
            outs.append(scatter_sum(_where(v >= 1e-10, v, 0.0), tsel, m))
        self._bond_candidates = sel
        self._bond_mask = mask
        return tuple(outs)

    # ---------------------------------------------------------- properties (167-178)
    def properties(self, x, charges, state):
        bo, total, nlp = state
        mass = self.a["mass"]
        center = np.sum(x * mass[:, None], axis=0) / np.sum(mass)
        n = self.n
        pair = np.bincount(self.i * n + self.j, weights=bo, minlength=n * n).reshape(n, n)
        counts = np.bincount(self.i, weights=(bo > BOND_GRAPH_CUT).astype(float), minlength=n).astype(int)
        return dict(
            bond_orders=pair,
            total_bond_orders=total,
            lone_pairs=nlp,
            bond_counts=counts,
            dipole=np.sum((x - center) * charges[:, None], axis=0),
        )

    # ---------------------------------------------------------- components (180-229)
    def components(self, x, fixed_charges=None, *, return_bond_state=False):
        a, ta, g = self.a, self.ta, self.g
        ti, tj = self.edges.ti, self.edges.tj
        vectors, r = self.edges.geometry(x)
        q, taper, shield = self.electrostatics(r, fixed_charges)
        bo, sigma, pi, pp = self.bond_orders(r)
        total = self.edges.atom_sum(bo)
        d, db = total - ta["valency"], total - ta["valency_boc"]
        de = total - ta["valency_e"]
        half = torch.trunc(de.detach() / 2)
        vlpex = de - 2 * half
        nlp = torch.exp(-g[15] * (2 + vlpex) ** 2) - half
        dlp = 0.5 * (ta["valency_e"] - ta["valency"]) - nlp
        dlpt = _where(a["mass"] > 21, 0, dlp)
        e = {k: torch.tensor(0.0, dtype=F64) for k in COMPONENTS}
        P = lambda k: _t(self.p(k))  # noqa: E731
        e["bond"] = -0.5 * torch.sum(
            P("De_s") * sigma * torch.exp(P("p_be1") * (1 - positive_power(sigma, P("p_be2"))))
            + P("De_p") * pi
            + P("De_pp") * pp
        )
        e["lone_pair"] = torch.sum(ta["p_lp2"] * dlp / (1 + torch.exp(-75 * dlp)))
        self.special_bond_corrections(e, bo, total, d)
        dfvl = _t((a["mass"] <= 21).astype(float))
        ov1 = self.edges.atom_sum(P("p_ovun1") * P("De_s") * bo)
        ov2 = self.edges.atom_sum((d[tj] - dfvl[ti] * dlpt[tj]) * (pi + pp))
        dc = d - dfvl * dlpt / (1 + g[32] * torch.exp(g[31] * ov2))
        over = ov1 * dc / (dc + ta["valency"] + 1e-8) / (1 + torch.exp(ta["p_ovun2"] * dc))
        under = (
            -ta["p_ovun5"]
            * (1 - torch.exp(g[6] * dc))
            / (1 + torch.exp(-ta["p_ovun2"] * dc))
            / (1 + g[8] * torch.exp(g[9] * ov2))
        )
        e["atom"] = torch.sum(over + under)
        if self.vdw_type in (1, 3):
            f13 = (r ** g[28] + (1 / P("gamma_w")) ** g[28]) ** (1 / g[28])
        else:
            f13 = r
        ev = torch.exp(0.5 * P("alpha") * (1 - f13 / P("r_vdW")))
        e["vdw"] = 0.5 * torch.sum(taper * P("D") * (ev**2 - 2 * ev))
        if self.vdw_type in (2, 3):
            e["vdw"] = e["vdw"] + 0.5 * torch.sum(taper * P("ecore") * torch.exp(P("acore") * (1 - r / P("rcore"))))
        e["coulomb"] = 0.5 * C_ELE * torch.sum(q[ti] * q[tj] * shield)
        e["qeq"] = SELF_CONVERSION * torch.sum(ta["chi"] * q + 0.5 * ta["eta"] * q * q)
        bv = bo.detach().numpy()
        active = np.flatnonzero(bv > THB_CUT)
        nbrs = _rows_of(active, self.i[active], self.n)
        self.angles(e, vectors, r, bo, pi, pp, total, d, db, nlp, vlpex, nbrs)
        self.torsions(e, vectors, r, bo, pi, db, nbrs)
        self.hydrogen_bonds(e, vectors, r, bo)
        comps = torch.stack([e[k] for k in COMPONENTS])
        if return_bond_state:
            return comps, q, tuple(v.detach().numpy().copy() for v in (bo, total, nlp))
        return comps, q

    # ------------------------------------------------ special corrections (231-246)
    def special_bond_corrections(self, e, bo, total, d):
        a, g = self.a, self.g
        i, j = self.i, self.j
        carbon = np.array([s.upper() == "C" for s in self.symbols])
        if g[5] > 0.001 and np.any(carbon):
            ti = torch.as_tensor(i)
            vov3 = bo - d[ti] - 0.040 * d[ti] ** 4
            mask = carbon[i] & carbon[j] & (vov3.detach().numpy() > 3)
            e["lone_pair"] = e["lone_pair"] + torch.sum(_where(mask, g[5] * (vov3 - 3) ** 2, 0))
        masses = a["mass"]
        co = ((masses[i] == 12.0) & (masses[j] == 15.999)) | ((masses[j] == 12.0) & (masses[i] == 15.999))
        allowed = co if int(g[37]) != 2 else np.ones(len(i), dtype=bool)
        sel = np.flatnonzero(allowed & (bo.detach().numpy() >= 1.0) & self.edges.half)
        if len(sel):
            si, sj = torch.as_tensor(i[sel]), torch.as_tensor(j[sel])
            b = bo[torch.as_tensor(sel)]
            stab = g[10] * torch.exp(-g[7] * (b - 2.5) ** 2) / (1 + 25 * torch.exp(g[4] * (d[si] + d[sj])))
            stab = stab * (torch.exp(-g[3] * (total[si] - b)) + torch.exp(-g[3] * (total[sj] - b)))
            e["bond"] = e["bond"] + torch.sum(stab)

    # ------------------------------------------------------ hydrogen bonds (248-276)
    def hydrogen_bond_items(self, bv, rv):
        roles = self.a["p_hbond"]
        hb_cut = min(HBOND_CUT, self.g[12])
        rp = self.edges.row_ptr
        donors_out, acceptors_out, params = [], [], []
        for h in np.flatnonzero(roles == 1):
            row = np.arange(rp[h], rp[h + 1])
            heavy = row[roles[self.j[row]] == 2]
            donors = heavy[bv[heavy] >= HBOND_THRESHOLD]
            acceptors = heavy[rv[heavy] <= hb_cut]
            if len(donors) == 0 or len(acceptors) == 0:
                continue
            dd = np.repeat(donors, len(acceptors))
            aa = np.tile(acceptors, len(donors))
            keep = dd != aa
            dd, aa = dd[keep], aa[keep]
            sh = self.symbols[h]
            for dn, ac in zip(dd, aa):
                prm = self.ff.hydrogen_bonds.get((self.symbols[self.j[dn]], sh, self.symbols[self.j[ac]]))
                if prm is not None and prm[0] > 0:
                    donors_out.append(dn)
                    acceptors_out.append(ac)
                    params.append(prm)
        return np.array(donors_out, dtype=np.int64), np.array(acceptors_out, dtype=np.int64), np.array(params).reshape(-1, 4)

    def hydrogen_bonds(self, e, vectors, r, bo):
        donor, acceptor, params = self.hydrogen_bond_items(bo.detach().numpy(), r.detach().numpy())
        self.hbond_items = (donor, acceptor)
        if len(donor) == 0:
            return
        r0, hb1, hb2, hb3 = (_t(c) for c in params.T)
        tdn, tac = torch.as_tensor(donor), torch.as_tensor(acceptor)
        dist = r[tac]
        cos = torch.sum(vectors[tdn] * vectors[tac], dim=1) / (r[tdn] * dist)
        sin4 = 0.25 * (1 - autograd_clip(cos, -1.0, 1.0)) ** 2
        e["hydrogen_bond"] = torch.sum(
            hb1 * (1 - torch.exp(-hb2 * bo[tdn])) * torch.exp(-hb3 * (r0 / dist + dist / r0 - 2)) * sin4
        )

    # --------------------------------------------------------------- angles (278-318)
    def angle_items(self, bv, nbrs):
        lefts, rights, params = [], [], []
        sym = self.symbols
        for c, row in enumerate(nbrs):
            k = len(row)
            if k < 2:
                continue
            for x in range(k - 1):
                for y in range(x + 1, k):
                    left, right = row[x], row[y]
                    if bv[left] * bv[right] <= THB_PRODUCT_CUT:
                        continue
                    key = (sym[self.j[left]], sym[c], sym[self.j[right]])
                    for prm in self.ff.angles.get(key, []):
                        if abs(prm[1]) > 0.001:
                            lefts.append(left)
                            rights.append(right)
                            params.append(prm)
        return np.array(lefts, dtype=np.int64), np.array(rights, dtype=np.int64), np.array(params).reshape(-1, 7)

    def angles(self, e, vectors, r, bo, pi, pp, total, d, db, nlp, vlpex, nbrs):
        a, ta, g = self.a, self.ta, self.g
        left, right, params = self.angle_items(bo.detach().numpy(), nbrs)
        self.angle_list = (left, right)
        if len(left) == 0:
            return
        ti_, tj_, tk_ = (torch.as_tensor(v) for v in (self.j[left], self.i[left], self.j[right]))
        tl, tr = torch.as_tensor(left), torch.as_tensor(right)
        theta00, v1, v2, coa1, v7, pen1, v4 = (_t(c) for c in params.T)
        prod = torch.exp(-self.edges.atom_sum(bo**8))
        sbo = self.edges.atom_sum(pi + pp) + (1 - prod) * (-db - g[33] * _where(vlpex.detach() >= 0, 0, nlp))
        sbo2 = torch.where(
            sbo <= 0,
            torch.zeros_like(sbo),
            torch.where(
                sbo <= 1,
                positive_power(sbo, g[16]),
                torch.where(sbo < 2, 2 - positive_power(2 - sbo, g[16]), torch.full_like(sbo, 2.0)),
            ),
        )
        theta0 = (180 - theta00 * (1 - torch.exp(-g[17] * (2 - sbo2[tj_])))) * (3.14159265 / 180)
        cos = torch.sum(vectors[tl] * vectors[tr], dim=1) / (r[tl] * r[tr])
        if self.lammps_forces:
            theta = lammps_angle(cos)
        else:
            theta = torch.arccos(autograd_clip(cos, -1 + 1e-14, 1 - 1e-14))
        bij, bjk = bo[tl] - THB_CUT, bo[tr] - THB_CUT
        pv3 = ta["p_val3"][tj_]
        f7 = (1 - torch.exp(-pv3 * bij**v4)) * (1 - torch.exp(-pv3 * bjk**v4))
        ex6, ex7 = torch.exp(g[14] * db[tj_]), torch.exp(-v7 * db[tj_])
        pv5 = ta["p_val5"][tj_]
        f8 = pv5 - (pv5 - 1) * (2 + ex6) / (1 + ex6 + ex7)
        exangle = torch.exp(-v2 * (theta0 - theta) ** 2)
        e["angle"] = torch.sum(f7 * f8 * v1 * torch.where(v1 >= 0, 1 - exangle, -exangle))
        ex3, ex4 = torch.exp(-g[20] * d[tj_]), torch.exp(g[21] * d[tj_])
        e["penalty"] = torch.sum(
            pen1 * (2 + ex3) / (1 + ex3 + ex4) * torch.exp(-g[19] * ((bij - 2) ** 2 + (bjk - 2) ** 2))
        )
        e["angle_conjugation"] = torch.sum(
            coa1
            / (1 + torch.exp(g[2] * (total[tj_] - ta["valency_val"][tj_])))
            * torch.exp(-g[38] * ((total[ti_] - bij) ** 2 + (total[tk_] - bjk) ** 2))
            * torch.exp(-g[30] * ((bij - 1.5) ** 2 + (bjk - 1.5) ** 2))
        )

    # ------------------------------------------------------------- torsions (320-370)
    def torsion_items(self, bv, nbrs):
        chains, params = [], []
        sym = self.symbols
        shifts = self.edges.shifts
        for middle in np.flatnonzero(self.edges.half & (bv > THB_CUT)):
            j, k = self.i[middle], self.j[middle]
            left_types = self.torsion_types.get((sym[j], sym[k]))
            if left_types is None:
                continue
            for left in nbrs[j]:
                if left == middle:
                    continue
                right_types = left_types.get(sym[self.j[left]])
                if right_types is None:
                    continue
                for right in nbrs[k]:
                    prm = right_types.get(sym[self.j[right]])
                    if prm is None:
                        continue
                    if right == self.edges.reverse[middle] or bv[left] * bv[middle] * bv[right] <= THB_CUT:
                        continue
                    i, l = self.j[left], self.j[right]
                    if i == l and np.array_equal(shifts[left], shifts[middle] + shifts[right]):
                        continue
                    chains.append((left, middle, right))
                    params.append(prm)
        return np.array(chains, dtype=np.int64).reshape(-1, 3), np.array(params).reshape(-1, 5)

    def torsions(self, e, vectors, r, bo, pi, db, nbrs):
        chains, params = self.torsion_items(bo.detach().numpy(), nbrs)
        self.torsion_list = chains
        if len(chains) == 0:
            return
        g = self.g
        left, middle, right = (torch.as_tensor(c) for c in chains.T)
        tj_, tk_ = torch.as_tensor(self.i[chains[:, 1]]), torch.as_tensor(self.j[chains[:, 1]])
        v1, v2, v3, tor1, cot1 = (_t(c) for c in params.T)
        u, v, w = vectors[left], vectors[middle], vectors[right]
        n1, n2 = torch.linalg.cross(u, v), torch.linalg.cross(w, v)
        area1, area2 = torch.sum(n1**2, dim=1), torch.sum(n2**2, dim=1)
        s1 = area1 / (r[left] * r[middle]) ** 2
        s2 = area2 / (r[middle] * r[right]) ** 2
        if bool(torch.any(s1.detach() < 1e-20)) or bool(torch.any(s2.detach() < 1e-20)):
            raise ValueError(
                "Collinear atoms in an active torsion: the dihedral derivative is undefined; perturb the geometry"
            )
        sinprod = torch.sqrt(s1 * s2)
        cw = torch.sum(n1 * n2, dim=1) / torch.sqrt(area1 * area2)
        b1, b2, b3 = bo[left] - THB_CUT, bo[middle] - THB_CUT, bo[right] - THB_CUT
        ds = db[tj_] + db[tk_]
        ex3, ex4 = torch.exp(-g[24] * ds), torch.exp(g[25] * ds)
        f11 = (2 + ex3) / (1 + ex3 + ex4)
        ex1 = torch.exp(tor1 * (2 - pi[middle] - f11) ** 2)
        fn10 = (1 - torch.exp(-g[23] * b1)) * (1 - torch.exp(-g[23] * b2)) * (1 - torch.exp(-g[23] * b3))
        cv = 0.5 * (v1 * (1 + cw) + v2 * ex1 * (2 - 2 * cw * cw) + v3 * (1 + 4 * cw**3 - 3 * cw))
        e["torsion"] = torch.sum(fn10 * sinprod * cv)
        fn12 = torch.exp(-g[27] * ((b1 - 1.5) ** 2 + (b2 - 1.5) ** 2 + (b3 - 1.5) ** 2))
        e["conjugation"] = torch.sum(cot1 * fn12 * (1 + (cw * cw - 1) * sinprod))


# ------------------------------------------------------------------- evaluation
def coincidence_check(x, lattice, pbc, periodic):
    """Reject atoms closer than 1e-6 A to any image of another atom (calculator.py:63-65)."""
    if len(x) < 2:
        return
    cut = 1e-6
    if periodic:
        i, j, S = image_neighbors(x, 1e-3, lattice, pbc)
    else:
        i, j, S = image_neighbors(x, 1e-3, np.eye(3), np.zeros(3, bool))
    if len(i) == 0:
        return
    off = shift_offsets(S, lattice) if periodic else np.zeros((len(i), 3))
    v = (x[j] - x[i]) + off
    dist = np.sqrt(np.sum(v * v, axis=1))
    if np.any((dist < cut) & (i != j)):
        raise ValueError("Coincident atoms are not supported")


def build_neighbors(x, cutoff, cell=None, pbc=None):
    lattice, flags, periodic = boundary(cell, pbc)
    return image_neighbors(x, cutoff, lattice if periodic else np.eye(3), flags)


def evaluate(ff, symbols, positions, *, total_charge=0, cell=None, pbc=None, full_derivative=False, neighbors=None):
    """Restatement of ``Calculator.evaluate`` (calculator.py:76-165), returning a dict."""
    symbols = tuple(symbols)
    x = np.array(positions, dtype=np.float64, copy=True)
    lattice, flags, periodic = boundary(cell, pbc)
    coincidence_check(x, lattice, flags, periodic)
    ff.validate_model(symbols)
    backend = "provided" if neighbors is not None else "replicated"
    if neighbors is None:
        neighbors = build_neighbors(x, ff.general[12], cell, pbc)
    model = Model(ff, symbols, neighbors, cell, pbc, total_charge=total_charge, lammps_forces=not full_derivative)
    fixed = None
    if not full_derivative:
        _, dist = model.edges.geometry(torch.as_tensor(x))
        fixed = model.electrostatics(dist)[0].numpy()
    xt = torch.tensor(x, dtype=F64, requires_grad=True)
    comps, q, state = model.components(xt, fixed, return_bond_state=True)
    total = comps.sum()
    (grad,) = torch.autograd.grad(total, xt, allow_unused=True)
    forces = -(grad.numpy() if grad is not None else np.zeros_like(x))
    comps_np = comps.detach().numpy()
    charges = q.detach().numpy().copy()
    if not all(np.isfinite(v).all() for v in (comps_np, charges, forces)):
        raise ValueError("Non-finite energy, charges, or forces; check the geometry")
    if abs(charges.sum() - total_charge) > 1e-8:
        raise ValueError("QEq charge constraint failed")
    props = model.properties(x, charges, state)
    out = dict(
        energy=float(comps_np.sum()),
        forces=forces,
        charges=charges,
        components=dict(zip(COMPONENTS, map(float, comps_np))),
        full_derivative=full_derivative,
        neighbor_backend=backend,
        model=model,
        **props,
    )
    return out


def fixed_charge_energy(model, x, charges):
    """Total energy at fixed charges (the objective whose gradient gives the forces)."""
    comps, _ = model.components(torch.as_tensor(np.asarray(x, dtype=float)), charges)
    return float(comps.detach().numpy().sum())
