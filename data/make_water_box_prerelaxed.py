"""Write data/water_box_12288_prerelaxed.npz: the benchmarked 12,288-atom water box.

Geometry: xreac_amd.workloads.water_box(side=16, seed=1) (4096 waters on a
jittered 16^3 lattice, 0.997 g/cm^3), then 200 steepest-descent steps
(workloads.steepest_descent: x += 0.002 F with each atom's step capped at
0.05 A) driven by the plain FP64 GPU forces at QEq tolerance 1e-10.  The
positions are stored so that the benchmark and the oracle fixture
(tests/golden/make_water_box_fixture.py) use the same coordinates bit for
bit.  Needs the GPU: python data/make_water_box_prerelaxed.py [out.npz]
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

STEPS = 200


def main(out=os.path.join(HERE, "water_box_12288_prerelaxed.npz")):
    import torch

    from xreac_amd import ForceField, workloads
    from xreac_amd.engine import Engine

    ff = ForceField.bundled("ffield.reax.HO.2015")
    eng = Engine(ff, 0)
    symbols, x0, cell = workloads.water_box(side=16, seed=1)
    types = torch.as_tensor(eng.type_ids(symbols), device=eng.device)

    def forces(y):
        r = eng.evaluate(torch.as_tensor(y, device=eng.device), types, cell=cell, pbc=(True,) * 3, tol=1e-10,
                         properties=False)
        return r["forces"].cpu().numpy()

    x = workloads.steepest_descent(forces, x0, steps=STEPS)
    np.savez_compressed(out, positions=x, cell=cell, steps=STEPS, seed=1, side=16)
    print(out, float(np.abs(x - x0).max()), flush=True)


if __name__ == "__main__":
    main(*sys.argv[1:])
