"""CG iteration blocks that own several chunks.

A system gets one CG chunk per 48 rows (per 24 rows below 8,192 atoms) and at
most XREAC_CG_MAX_BLOCKS iteration blocks (256 by default, 512 below 8,192
atoms), so systems above 12,288 atoms give some blocks several chunks, which
they sweep one after the other.  The cap is read once per process, so a child
process with a cap of 4 evaluates the 3150-atom CHO mixture (small block
shape: 132 chunks, 33 per block) and an 8,232-atom water box (large block
shape: 172 chunks, 43 per block), and their energy, forces and charges must
agree with this process's evaluation (one chunk per block).  The dot partials
are grouped differently, so the two solves differ in their last bits only;
both are converged to tolerance 1e-12.
"""

import os
import subprocess
import sys

import numpy as np
import pytest

from xreac_amd import ForceField, workloads

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
import numpy as np
import torch
from xreac_amd import ForceField, workloads
from xreac_amd.engine import Engine
if sys.argv[3] == "cho":
    s, x, cell = workloads.cho_mixture()
    eng = Engine(ForceField.bundled("ffield.reax.CHO.2008"))
else:
    s, x, cell = workloads.water_box(side=14, seed=7)
    eng = Engine(ForceField.bundled("ffield.reax.HO.2015"))
xd = torch.as_tensor(np.asarray(x, dtype=np.float64), device=eng.device).contiguous()
td = torch.as_tensor(eng.type_ids(s), device=eng.device)
r = eng.evaluate(xd, td, cell=cell, pbc=(True,) * 3, tol=1e-12, precision=sys.argv[1])
np.savez(sys.argv[2], e=float(r["components"].sum()), f=r["forces"].cpu().numpy(), q=r["charges"].cpu().numpy())
"""

_ENG = {}


SYSTEMS = {"cho": "ffield.reax.CHO.2008", "box8232": "ffield.reax.HO.2015"}


def _evaluate(system, precision):
    import torch
    from xreac_amd.engine import Engine

    if system not in _ENG:
        _ENG[system] = Engine(ForceField.bundled(SYSTEMS[system]))
    eng = _ENG[system]
    if system == "cho":
        s, x, cell = workloads.cho_mixture()
    else:
        s, x, cell = workloads.water_box(side=14, seed=7)
    xd = torch.as_tensor(np.asarray(x, dtype=np.float64), device=eng.device).contiguous()
    td = torch.as_tensor(eng.type_ids(s), device=eng.device)
    return eng.evaluate(xd, td, cell=cell, pbc=(True,) * 3, tol=1e-12, precision=precision)


@pytest.mark.parametrize("system", list(SYSTEMS))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_blocks_with_several_chunks_agree(tmp_path, precision, system):
    out = tmp_path / "child.npz"
    env = dict(os.environ, XREAC_CG_MAX_BLOCKS="4")
    print(f"[cg-chunks] child process, XREAC_CG_MAX_BLOCKS=4, {system}, {precision}", flush=True)
    subprocess.run([sys.executable, "-c", CHILD, precision, str(out), system], env=env, cwd=ROOT, check=True,
                   timeout=110)
    child = np.load(out)
    print(f"[cg-chunks] this process, default cap, {system}, {precision}", flush=True)
    r = _evaluate(system, precision)
    e = float(r["components"].sum())
    assert abs(float(child["e"]) - e) <= 1e-10 * abs(e)
    np.testing.assert_allclose(child["f"], r["forces"].cpu().numpy(), rtol=0, atol=1e-7)
    np.testing.assert_allclose(child["q"], r["charges"].cpu().numpy(), rtol=0, atol=1e-9)
