"""Evaluations whose first attempt overflows its bond buffers.

An evaluation sizes its edge and bond buffers from the previous call's counts
per atom; when the counts outgrow them, the attempt is repeated with exact
sizes.  The bond list is built on the second stream beside the QEq solve, so
an overflow there is found while or after the solve runs.  In a child process
XREAC_BOND_CAPACITY_SCALE (read once per process) shrinks that sized bond
capacity so that every such first attempt overflows; the child makes the same
sequence of calls as this process, and every result must be bitwise the same,
warm starts included: the repeat starts from the caller's guess, not from the
first attempt's solution, and the per-bond kernels stay inside the buffers of
the attempt that overflowed.  No reference counterpart (the reference sizes
nothing ahead).
"""

import os
import subprocess
import sys

import numpy as np
import pytest

from xreac_amd import ForceField, workloads

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HO = "ffield.reax.HO.2015"

CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, ".")
from tests.test_gpu_overflow import sequence
np.savez(sys.argv[1], **sequence())
"""


def sequence():
    """A four-system batch and a lone system, each evaluated cold, then twice warm from moved positions."""
    import torch

    from xreac_amd.engine import Engine

    eng = Engine(ForceField.bundled(HO))
    s1, x1 = workloads.water_cluster_of(40, 3)
    s2, x2, c2 = workloads.water_cell(32, seed=4)
    s3, x3 = workloads.water_cluster_of(25, 5)
    s4, x4, q4 = workloads.ion_cluster(20, -1, 6)
    s5, x5 = workloads.water_cluster()
    cases = {
        "batch": ([s1, s2, s3, s4], [x1, x2, x3, x4], [np.eye(3), np.asarray(c2), np.eye(3), np.eye(3)],
                  [[0, 0, 0], [1, 1, 1], [0, 0, 0], [0, 0, 0]], [0.0, 0.0, 0.0, q4]),
        "single": ([s5], [x5], [np.eye(3)], [[0, 0, 0]], [0.0]),
    }
    out = {}
    rng = np.random.default_rng(11)
    for name, (syms, xs, cells, pbcs, charges) in cases.items():
        sym = [a for s in syms for a in s]
        offsets = np.array([0] + list(np.cumsum([len(s) for s in syms])), dtype=np.int32)
        x0 = np.vstack(xs)
        t = torch.as_tensor(eng.type_ids(sym), device=eng.device)
        o = eng.batch_outputs(len(sym), qeq_state=True)
        for call in range(3):
            x = x0 + (0.02 * call) * rng.standard_normal(x0.shape)
            xd = torch.as_tensor(x, device=eng.device).contiguous()
            r = eng.evaluate_batch(xd, t, offsets, np.stack(cells), np.array(pbcs, dtype=np.int32),
                                   np.array(charges), tol=1e-10, warm_start=call > 0, out=o)
            assert all(e is None for e in r["errors"]), r["errors"]
            key = f"{name}_{call}"
            out[key + "_E"] = r["components"]
            out[key + "_F"] = r["forces"].cpu().numpy()
            out[key + "_Q"] = r["charges"].cpu().numpy()
            out[key + "_S"] = r["qeq_state"].cpu().numpy()
            out[key + "_it"] = r["qeq_iterations"]
            out[key + "_bonds"] = np.array(r["num_bonds"])
    return out


def test_bond_overflow_repeat_is_bitwise(tmp_path):
    path = tmp_path / "child.npz"
    env = dict(os.environ, XREAC_BOND_CAPACITY_SCALE="0.3")
    print("[overflow] child process, XREAC_BOND_CAPACITY_SCALE=0.3", flush=True)
    subprocess.run([sys.executable, "-c", CHILD, str(path)], env=env, cwd=ROOT, check=True, timeout=110)
    child = np.load(path)
    print("[overflow] this process, default capacity", flush=True)
    mine = sequence()
    assert sorted(child.files) == sorted(mine)
    for key in mine:
        np.testing.assert_array_equal(child[key], mine[key], err_msg=key)
