"""Write tests/golden/water_box_12288.npz: CPU-oracle results for the benchmarked 12,288-atom box.

The geometry is the pre-relaxed box the benchmark times
(data/water_box_12288_prerelaxed.npz: workloads.water_box(side=16, seed=1)
after 200 steepest-descent steps), so only results are stored: energy,
components, the neighbor-list and active-bond-set digests, and
charges/forces/bond properties on 512 seeded atoms.  Run (CPU, about two
minutes on eight cores): python tests/golden/make_water_box_fixture.py
"""

import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import reaxff as orc  # noqa: E402
from xreac_amd import ForceField, workloads  # noqa: E402


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, dtype=np.int64).tobytes())
    return h.hexdigest()


def main():
    ff = ForceField.bundled("ffield.reax.HO.2015")
    s, x, cell = workloads.prerelaxed_water_box()
    r = orc.evaluate(ff, s, x, cell=cell, pbc=True)
    model = r["model"]
    active = model._bond_candidates[model._bond_mask]
    subset = np.sort(np.random.default_rng(12288).choice(len(s), 512, replace=False))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "water_box_12288.npz")
    np.savez_compressed(
        out,
        energy=r["energy"],
        components=np.array(list(r["components"].values())),
        num_edges=len(model.i),
        neighbor_sha256=digest(model.i, model.j, model.edges.shifts),
        num_bonds=len(active),
        bond_sha256=digest(model.i[active], model.j[active], model.edges.shifts[active]),
        subset=subset,
        charges=r["charges"][subset],
        forces=r["forces"][subset],
        total_bond_orders=r["total_bond_orders"][subset],
        lone_pairs=r["lone_pairs"][subset],
        bond_counts=r["bond_counts"][subset],
        charge_sum=r["charges"].sum(),
        force_sum=r["forces"].sum(axis=0),
        positions_sha256=hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest(),
    )
    print(out, r["energy"], len(model.i), len(active))


if __name__ == "__main__":
    main()
