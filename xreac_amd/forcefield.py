"""ReaxFF parameter files in the standard 39-global, four-line-atom layout.

Behaviour follows the reference reader ``src/xreac/forcefield.py`` (hsulab/xreac
0.10.0): the same section grammar, the same derived tables and mixing rules, the
same error classes.  Differences in structure are deliberate; the observable
tables (``general``, ``atoms``, ``pairs``, ``angles``, ``torsions``,
``hydrogen_bonds``) carry the same keys and values.

Conventions kept from the reference (forcefield.py line numbers):

* ``eta`` is stored doubled (178) and ``p_hbond`` as an int (179); atoms lighter
  than 21 amu use ``valency_boc`` as ``valency_val`` (180-181).
* Every unordered type pair gets nonbonded/bond-order mixing, with or without an
  explicit bond record (211-244): arithmetic means for ``r_s``/``r_p``/``r_pp``,
  geometric means elsewhere, ``r_vdW`` doubled, ``gamma = (g_i g_j)^-1.5``;
  a sign mismatch is an error except for ``p_boc3/4/5``, which become NaN and
  are validated where an active correction would use them.
* Off-diagonal overrides apply only to positive values; ``r_vdW`` overrides are
  doubled (245-254).
* Angle rows accumulate per key in both orientations (255-262).
* Torsion rows: explicit keys always win over ``0-j-k-0`` wildcards, later
  wildcard rows replace earlier ones (263-286); only the first five values are
  kept (272).
* The hydrogen-bond block may be omitted (287-291).
"""

from __future__ import annotations

import hashlib
import pickle
import warnings
from dataclasses import dataclass, field
from itertools import combinations_with_replacement
from pathlib import Path

import numpy as np

# Column names of the 32 atom parameters and the 16 bond parameters, in file
# order.  These are the public dictionary keys (forcefield.py:16-24).
ATOM_NAMES = tuple(
    """r_s valency mass r_vdw epsilon gamma r_pi valency_e
    alpha gamma_w valency_boc p_ovun5 unused12 chi eta p_hbond
    r_pi_pi p_lp2 heat b_o_131 b_o_132 b_o_133 bcut_acks2 unused23
    p_ovun2 p_val3 unused26 valency_val p_val5 rcore ecore acore""".split()
)
BOND_NAMES = tuple(
    "De_s De_p De_pp p_be1 p_bo5 v13cor p_bo6 p_ovun1 p_be2 p_bo3 p_bo4 unused11 p_bo1 p_bo2 ovc unused15".split()
)
SECTION_NAMES = ("general", "atoms", "bonds", "off_diagonal", "angles", "torsions", "hydrogen_bonds")

# Pair quantities mixed from atom columns: (pair key, atom column).
_ARITHMETIC = (("r_s", "r_s"), ("r_p", "r_pi"), ("r_pp", "r_pi_pi"))
_GEOMETRIC = (
    ("p_boc3", "b_o_132"),
    ("p_boc4", "b_o_131"),
    ("p_boc5", "b_o_133"),
    ("D", "epsilon"),
    ("alpha", "alpha"),
    ("gamma_w", "gamma_w"),
    ("r_vdW", "r_vdw"),
    ("rcore", "rcore"),
    ("ecore", "ecore"),
    ("acore", "acore"),
)
_DEFERRED_NAN = frozenset({"p_boc3", "p_boc4", "p_boc5"})
_OVERRIDES = ("D", "r_vdW", "alpha", "r_s", "r_p", "r_pp")

DATA_DIR = Path(__file__).resolve().parent.parent / "data"


class _Lines:
    """Iterator over significant tokenised lines (comments after ! or # dropped)."""

    def __init__(self, lines):
        self._lines = lines
        self._pos = 0

    def raw(self):
        """Consume one physical line without interpreting it (header rows)."""
        if self._pos >= len(self._lines):
            raise StopIteration
        self._pos += 1

    def tokens(self):
        while self._pos < len(self._lines):
            text = self._lines[self._pos]
            self._pos += 1
            toks = text.split("!")[0].split("#")[0].split()
            if toks:
                return toks
        raise StopIteration

    def count(self):
        n = int(self.tokens()[0])
        if n < 0:
            raise ValueError("Negative section count")
        return n

    def trailing(self):
        rest = self._lines[self._pos :]
        self._pos = len(self._lines)
        return any(t.split("!")[0].split("#")[0].strip() for t in rest)


def _numbers(tokens, expected):
    if len(tokens) != expected:
        raise ValueError(f"Expected {expected} parameters, received {len(tokens)}")
    values = np.array([float(t.replace("D", "E").replace("d", "e")) for t in tokens], dtype=np.float64)
    if not np.isfinite(values).all():
        raise ValueError("Non-finite force-field parameters")
    return values


@dataclass(frozen=True)
class ForceField:
    """Parsed parameter tables (interface of ``src/xreac/forcefield.py:27-341``)."""

    path: Path
    checksum: str
    citation: str
    general: np.ndarray
    atoms: dict
    pairs: dict
    angles: dict
    torsions: dict
    hydrogen_bonds: dict
    _sections: tuple = field(default=(), repr=False, compare=False)
    _parsed_state: bytes = field(default=b"", repr=False, compare=False)

    # ------------------------------------------------------------------ state
    def _state(self):
        # Digest of the in-memory tables; detects edits before export.
        blob = pickle.dumps((self.general, self.atoms, self.pairs, self.angles, self.torsions, self.hydrogen_bonds))
        return hashlib.sha256(blob).digest()

    @property
    def section_counts(self):
        """Explicit record counts per section, before mixing or wildcard expansion."""
        if not self._sections:
            raise ValueError("Source parameter records are unavailable; load with from_file")
        return {name: len(records) for name, records in zip(SECTION_NAMES, self._sections)}

    @property
    def elements(self):
        """Atom labels in file order (dummy types included)."""
        return tuple(self.atoms)

    @property
    def vdw_type(self):
        """LAMMPS vdW variant: 1 shielded, 2 inner wall, 3 both (forcefield.py:102-118)."""
        kinds = []
        for a in self.atoms.values():
            shielded = a["gamma_w"] > 0.5
            wall = a["rcore"] > 0.01 and a["acore"] > 0.01
            kinds.append(3 if (shielded and wall) else 1 if shielded else 2 if wall else 0)
        if not kinds or 0 in kinds:
            raise ValueError("Inconsistent or unsupported van der Waals method across atom types")
        if len(set(kinds)) > 1:
            warnings.warn(
                "Mixed van der Waals methods; using the first atom type's method, as LAMMPS does",
                RuntimeWarning,
                stacklevel=2,
            )
        return kinds[0]

    # ----------------------------------------------------------------- export
    def to_file(self, path):
        """Write the retained source records (lossless for unmodified objects)."""
        if not self._sections or self._state() != self._parsed_state:
            raise ValueError("Writing requires an unmodified ForceField loaded by from_file")

        def fmt(v):
            return str(v) if isinstance(v, (str, int)) else format(v, ".17g")

        out = [self.citation]
        for index, (name, records) in enumerate(zip(SECTION_NAMES, self._sections)):
            out.append(f"{len(records)} ! {name}")
            out.extend(["! parameter columns"] * {1: 3, 2: 1}.get(index, 0))
            for record in records:
                toks = [fmt(v) for v in record]
                if index == 1:  # atom: name + 8, then three rows of 8
                    out.append(" ".join(toks[:9]))
                    out.extend(" ".join(toks[k : k + 8]) for k in (9, 17, 25))
                elif index == 2:  # bond: two indices + 8, then 8
                    out.append(" ".join(toks[:10]))
                    out.append(" ".join(toks[10:]))
                else:
                    out.append(" ".join(toks))
        Path(path).write_text("\n".join(out) + "\n")

    # ---------------------------------------------------------------- loading
    @classmethod
    def bundled(cls, name):
        """Load a parameter file shipped in this repository's ``data/`` directory."""
        if not isinstance(name, str) or name in ("", ".", "..") or Path(name).name != name:
            raise ValueError("A bundled parameter name must be a plain filename")
        return cls.from_file(DATA_DIR / name)

    @classmethod
    def from_file(cls, path):
        """Parse a ReaxFF file; malformed content raises ValueError, missing files OSError."""
        path = Path(path).resolve()
        blob = path.read_bytes()
        text_lines = blob.decode().splitlines()
        if not text_lines:
            raise ValueError(f"Empty ReaxFF parameter file: {path}")
        citation, body = text_lines[0], _Lines(text_lines[1:])
        try:
            tables = _parse_body(body)
        except (StopIteration, IndexError, ValueError) as exc:
            raise ValueError(f"Invalid or unsupported ReaxFF file {path}: {exc}") from exc
        general, atoms, pairs, angles, torsions, hbonds, sections = tables
        ff = cls(
            path,
            hashlib.sha256(blob).hexdigest(),
            citation,
            general,
            atoms,
            pairs,
            angles,
            torsions,
            hbonds,
            tuple(tuple(s) for s in sections),
        )
        object.__setattr__(ff, "_parsed_state", ff._state())
        return ff

    # ------------------------------------------------------------- validation
    def validate_model(self, elements=None):
        """Reject parameter sets outside the implemented equations (forcefield.py:317-341)."""
        self.vdw_type
        g = self.general
        if g[11] != 0 or g[12] <= 0:
            raise ValueError("Only zero lower taper radius and a positive upper cutoff are supported")
        if g[1] <= 0 or g[28] <= 0 or g[29] <= 0:
            raise ValueError("Invalid bond-order correction, vdW exponent, or bond-order cutoff")
        if elements is None:
            return
        present = set(elements)
        missing = present - self.atoms.keys()
        if missing:
            raise ValueError(f"Atom types absent from the force field: {', '.join(sorted(missing))}")
        for s in present:
            a = self.atoms[s]
            if a["eta"] <= 0 or a["gamma"] <= 0:
                raise ValueError("Invalid QEq hardness or shielding")
            if a["mass"] <= 0 or a["valency"] < 0:
                raise ValueError("Positive atomic masses and nonnegative valencies are required")
            for t in present:
                p = self.pairs[s, t]
                if p["r_vdW"] <= 0:
                    raise ValueError(f"Invalid van der Waals radius for {s}/{t}")
                for atom_key, pair_key in (("r_s", "r_s"), ("r_pi", "r_p"), ("r_pi_pi", "r_pp")):
                    if a[atom_key] > 0 and self.atoms[t][atom_key] > 0 and p[pair_key] <= 0:
                        raise ValueError(f"Invalid {pair_key} for {s}/{t}")


def _parse_body(lines: _Lines):
    sections = [[] for _ in SECTION_NAMES]

    n_general = lines.count()
    if n_general != 39:
        raise ValueError("Only the standard 39-global-parameter format is supported")
    general = np.array([_numbers(lines.tokens()[:1], 1)[0] for _ in range(n_general)])
    sections[0] = [(float(v),) for v in general]

    # ---- atoms: count line, three header lines, then 4 rows per atom
    n_atoms = lines.count()
    if n_atoms == 0:
        raise ValueError("A force field must define at least one atom type")
    for _ in range(3):
        lines.raw()
    atoms, names = {}, []
    for _ in range(n_atoms):
        first = lines.tokens()
        label = first[0]
        if any(label.casefold() == other.casefold() for other in names):
            raise ValueError(f"Duplicate element {label}")
        vals = _numbers(first[1:] + lines.tokens() + lines.tokens() + lines.tokens(), 32)
        sections[1].append((label, *vals.tolist()))
        rec = {k: float(v) for k, v in zip(ATOM_NAMES, vals)}
        rec["eta"] *= 2
        rec["p_hbond"] = int(rec["p_hbond"])
        if rec["mass"] < 21:
            rec["valency_val"] = rec["valency_boc"]
        atoms[label] = rec
        names.append(label)

    def labels(tokens, wildcard=False):
        out = []
        for tok in tokens:
            k = int(tok)
            if wildcard and k == 0:
                out.append("*")
            elif 1 <= k <= len(names):
                out.append(names[k - 1])
            else:
                raise ValueError("Invalid element index")
        return tuple(out)

    # ---- bonds: count line + one header line, two rows per record
    n_bonds = lines.count()
    lines.raw()
    pairs = {}
    for _ in range(n_bonds):
        first = lines.tokens()
        key = labels(first[:2])
        second = lines.tokens()
        if len(second) == 7:  # the unused 16th value may be omitted
            second = second + ["0"]
        vals = _numbers(first[2:] + second, 16)
        sections[2].append((int(first[0]), int(first[1]), *vals.tolist()))
        rec = {k: float(v) for k, v in zip(BOND_NAMES, vals)}
        pairs[key] = rec
        pairs[key[::-1]] = rec

    for s, t in combinations_with_replacement(names, 2):
        rec = pairs.get((s, t)) or {k: 0.0 for k in BOND_NAMES}
        a, b = atoms[s], atoms[t]
        for out, col in _ARITHMETIC:
            rec[out] = (a[col] + b[col]) / 2
        for out, col in _GEOMETRIC:
            prod = a[col] * b[col]
            if prod < 0:
                if out not in _DEFERRED_NAN:
                    raise ValueError(f"Invalid mixing parameters for {s}/{t}: {col}")
                rec[out] = float("nan")
            else:
                rec[out] = float(np.sqrt(prod))
        rec["r_vdW"] *= 2
        if a["gamma"] * b["gamma"] <= 0:
            raise ValueError("Charge shielding parameters must be positive")
        rec["gamma"] = float((a["gamma"] * b["gamma"]) ** -1.5)
        pairs[(s, t)] = rec
        pairs[(t, s)] = rec

    # ---- off-diagonal overrides
    for _ in range(lines.count()):
        toks = lines.tokens()
        key = labels(toks[:2])
        vals = _numbers(toks[2:], 6)
        sections[3].append((int(toks[0]), int(toks[1]), *vals.tolist()))
        if key not in pairs:
            raise ValueError("Off-diagonal entry without bond parameters")
        for name, v in zip(_OVERRIDES, vals):
            if v > 0:
                pairs[key][name] = float(v) * (2 if name == "r_vdW" else 1)

    # ---- angles (multi-row, both orientations, palindromes once)
    angles = {}
    for _ in range(lines.count()):
        toks = lines.tokens()
        key = labels(toks[:3])
        vals = _numbers(toks[3:], 7)
        sections[4].append((*map(int, toks[:3]), *vals.tolist()))
        for k in {key, key[::-1]}:
            angles.setdefault(k, []).append(vals)

    # ---- torsions with 0-j-k-0 wildcards
    torsions, explicit = {}, set()
    for _ in range(lines.count()):
        toks = lines.tokens()
        key = labels(toks[:4], wildcard=True)
        n_vals = len(toks) - 4
        if n_vals not in (5, 6, 7):
            raise ValueError("Expected 5 to 7 torsion parameters")
        full = _numbers(toks[4:], n_vals)
        sections[5].append((*map(int, toks[:4]), *full.tolist()))
        vals = full[:5]
        if "*" not in key:
            torsions[key] = vals
            torsions[key[::-1]] = vals
            explicit.update((key, key[::-1]))
            continue
        if key[0] != "*" or key[3] != "*" or "*" in key[1:3]:
            raise ValueError("Only terminal 0-i-j-0 torsion wildcards are supported")
        for s in names:
            for t in names:
                for k in ((s, key[1], key[2], t), (t, key[2], key[1], s)):
                    if k not in explicit:
                        torsions[k] = vals

    # ---- hydrogen bonds (block optional)
    hbonds = {}
    try:
        n_hb = lines.count()
    except StopIteration:
        n_hb = 0
    for _ in range(n_hb):
        toks = lines.tokens()
        key = labels(toks[:3])
        vals = _numbers(toks[3:], 4)
        hbonds[key] = vals
        sections[6].append((*map(int, toks[:3]), *vals.tolist()))
    if lines.trailing():
        raise ValueError("Unsupported trailing parameter section")
    if not np.isfinite(general).all():
        raise ValueError("Non-finite general parameters")
    return general, atoms, pairs, angles, torsions, hbonds, sections
