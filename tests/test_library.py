"""CPU checks of the C-ABI library and the host-side table packing (no GPU calls)."""

import ctypes
import os
import re

import numpy as np
import pytest

from xreac_amd import ForceField, _lib
from xreac_amd.tables import build_tables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "xreac.h")).read()
    return sorted(set(re.findall(r"\b(xreac_[a-z_]+)\s*\(", text)))


def test_library_loads_and_exports_header_symbols():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libxreac_hip.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared_symbols()
    assert declared, "no declarations parsed from include/xreac.h"
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.EXPORTED)
    _lib.load()


def test_engine_refuses_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from xreac_amd import Calculator

    calc = Calculator(ForceField.bundled("ffield.reax.HO.2015"))
    with pytest.raises(RuntimeError, match="GPU"):
        calc.evaluate(["O", "H", "H"], [[0, 0, 0], [0.97, 0, 0], [-0.243, 0.94, 0]])


@pytest.mark.parametrize("name", ["ffield.reax.HO.2015", "ffield.reax.CHO.2008", "ffield.reax.TiOH.Kim2013"])
def test_tables_match_force_field(name):
    ff = ForceField.bundled(name)
    t = build_tables(ff)
    nt = len(ff.elements)
    assert t.atom.shape == (nt, _lib_atom_width())
    for ti, s in enumerate(ff.elements):
        assert t.atom[ti, 9] == ff.atoms[s]["eta"]  # doubled eta
        for tj, u in enumerate(ff.elements):
            assert t.pair[ti * nt + tj, 27] == ff.pairs[s, u]["gamma"]
    # every angle row with |v1| > 0.001 appears under its key
    for (a, b, c), rows in ff.angles.items():
        key = (t.index[a] * nt + t.index[b]) * nt + t.index[c]
        kept = [r for r in rows if abs(r[1]) > 0.001]
        got = t.angle_rows[t.angle_offsets[key] : t.angle_offsets[key + 1]]
        np.testing.assert_array_equal(got, np.array(kept).reshape(-1, 7))
    for (a, b, c, d), prm in ff.torsions.items():
        key = ((t.index[a] * nt + t.index[b]) * nt + t.index[c]) * nt + t.index[d]
        np.testing.assert_array_equal(t.torsion[key], prm[:5])
        assert t.torsion_on[key] == int(np.any(np.asarray(prm)[[0, 1, 2, 4]] != 0))


def _lib_atom_width():
    text = open(os.path.join(ROOT, "include", "xreac.h")).read()
    return int(re.search(r"XREAC_ATOM_PARAMS\s*=\s*(\d+)", text).group(1))
