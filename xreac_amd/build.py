"""Build ``libxreac_hip.so`` in-tree with hipcc for gfx950.

``python -m xreac_amd.build`` or ``__graft_entry__.build()``.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = (
    "scan.hip",
    "neighbor.hip",
    "qeq.hip",
    "bondorder.hip",
    "manybody.hip",
    "pairs.hip",
    "integrate.hip",
    "engine.hip",
)
FLAGS = (
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # Parity with the reference's rounding: no fused multiply-add contraction.
    "-ffp-contract=off",
    # Packed FP32 instructions are disabled on this target.
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
)


def build(verbose=False):
    out = os.path.join(HERE, "libxreac_hip.so")
    tmp = out + ".tmp"
    cmd = ["hipcc", *FLAGS, "-o", tmp, *[os.path.join(HERE, "csrc", s) for s in SOURCES]]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    log = "\n".join(line for line in (proc.stdout + proc.stderr).splitlines() if noise not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{log}")
    if verbose and log.strip():
        print(log, file=sys.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
