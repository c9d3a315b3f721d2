"""Two contexts evaluating at once from two host threads.

The serving loop runs one engine per force field, each in its own thread; the
library lets one evaluation at a time run per device (evaluations whose kernels
overlapped gave rare run-to-run differences).  Each thread here steps its own
batch through NVE steps (velocity Verlet around a warm-started batch
evaluation) while the other does the same; every result must equal, bitwise,
the same steps run with the other thread idle.
"""

import threading

import numpy as np
import pytest

from xreac_amd import ForceField, workloads

pytestmark = pytest.mark.gpu


def _batch(ff_name, seed):
    rng = np.random.default_rng(seed)
    family = 0.1 if ff_name.endswith("CHO.2008") else 0.5  # the stream's CHO and water-cluster families
    systems = []
    for k in range(12):
        spec = dict(rid=k, atoms=int(rng.integers(60, 220)), family=family, periodic=False, seed=seed * 100 + k,
                    md_steps=0, dt=0.25, temperature=300.0, sign=1)
        r = workloads.build_request(spec)
        assert r["force_field"] == ff_name
        systems.append((list(r["symbols"]), np.asarray(r["positions"], dtype=np.float64)))
    return systems


def _run(ff_name, seed, steps, barrier=None):
    import torch

    from xreac_amd.engine import Engine

    ff = ForceField.bundled(ff_name)
    eng = Engine(ff, stream=torch.cuda.Stream(0))
    systems = _batch(ff_name, seed)
    sym = [a for s, _ in systems for a in s]
    offsets = np.array([0] + list(np.cumsum([len(s) for s, _ in systems])), dtype=np.int32)
    masses = np.array([ff.atoms[a]["mass"] for a in sym])
    v0 = workloads.maxwell_boltzmann(masses, 300.0, seed=seed)
    ns, n = len(systems), len(sym)
    out = []
    with torch.cuda.stream(eng.stream):
        dev = eng.device
        f64 = dict(dtype=torch.float64, device=dev)
        x = torch.as_tensor(np.vstack([xx for _, xx in systems]), **f64).contiguous()
        v = torch.as_tensor(v0, **f64).contiguous()
        t = torch.as_tensor(eng.type_ids(sym), device=dev)
        o = eng.batch_outputs(n, qeq_state=True)
        off = torch.as_tensor(offsets, device=dev)
        dt = torch.full((ns,), 0.25, **f64)
        if barrier is not None:
            barrier.wait()
        r = eng.evaluate_batch(x, t, offsets, tol=1e-10, warm_start=True, out=o)
        for _ in range(steps):
            eng.verlet(off, t, dt, x, v, o["forces"], 0)
            r = eng.evaluate_batch(x, t, offsets, tol=1e-10, warm_start=True, out=o)
            eng.verlet(off, t, dt, x, v, o["forces"], 1)
            out.append(r["components"].copy())
        torch.cuda.synchronize()
        return out, x.cpu().numpy(), o["charges"].cpu().numpy()


def test_two_contexts_at_once_match_each_alone():
    jobs = [("ffield.reax.HO.2015", 5), ("ffield.reax.CHO.2008", 6)]
    alone = [_run(name, seed, 15) for name, seed in jobs]
    print("[concurrent] alone done", flush=True)
    together = [None, None]
    barrier = threading.Barrier(2)

    def work(k):
        together[k] = _run(*jobs[k], 15, barrier)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=150)
    print("[concurrent] together done", flush=True)
    for k in range(2):
        assert together[k] is not None, jobs[k]
        comps_a, x_a, q_a = alone[k]
        comps_b, x_b, q_b = together[k]
        for step, (a, b) in enumerate(zip(comps_a, comps_b)):
            np.testing.assert_array_equal(a, b, err_msg=f"{jobs[k][0]} step {step}")
        np.testing.assert_array_equal(x_a, x_b)
        np.testing.assert_array_equal(q_a, q_b)
