#!/usr/bin/env python
"""Benchmark: ReaxFF energy + forces + QEq evaluations on one MI355X.

Workload (default): the 4096-water periodic box, 12,288 atoms, ffield.reax.HO.2015,
plain FP64, QEq relative residual 1e-10.  Geometry: 16^3 jittered lattice with
random orientations (seed 1), then 200 steepest-descent steps on the GPU
(x += 0.002 F, per-atom step capped at 0.05 A) before timing.

A step is one complete evaluation (neighbor build, QEq, all energy terms,
fixed-charge forces, bond properties) with positions already resident in HBM.
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TIMED_EVERY = 8  # timed region: events on one dispatch in 8 of the dominant kernel


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_bytes(name, n, m, nb):
    """Algorithmic HBM bytes per launch of a timed kernel class.

    cg_iteration: the implicit QEq matrix read once in half storage (value 8 B
    + column 4 B per unordered pair, P = M/2) plus 436 B per atom: the eight
    recurrence vectors read and written (x, r, u, w, z, q, s, p as FP64
    pairs: 256 B), m read once and written once (32 B), the block-Jacobi row
    (16 doubles: 128 B), eta, block placement and row pointer (20 B).
    """
    P = m / 2
    if name == "cg_iteration":
        return 12 * P + 436 * n
    if name == "nonbonded":
        return 16 * P + 56 * n  # half list (j + image code) + positions, types, charges, forces
    if name == "edge":
        return 8 * P + 12 * P + 40 * n  # half list read, H write, positions
    if name == "neighbor_fill":
        return 16 * P + 32 * n
    return None


def build_workload(args, eng, torch):
    from xreac_amd import workloads

    symbols, x, cell = workloads.water_box(side=args.side, seed=1)
    types = torch.as_tensor(eng.type_ids(symbols), device=eng.device)
    pbc = (True, True, True)
    cache = args.geometry_cache
    if cache and os.path.exists(cache):
        data = np.load(cache)
        if data["x"].shape == x.shape and int(data["prerelax"]) == args.prerelax:
            log(f"[bench] loaded pre-relaxed geometry from {cache}")
            return symbols, data["x"], cell, types, pbc

    def forces(y):
        xd = torch.as_tensor(y, device=eng.device)
        out = eng.evaluate(xd, types, cell=cell, pbc=pbc, tol=args.qeq_tol, properties=False)
        return out["forces"].cpu().numpy()

    t0 = time.time()
    x = workloads.steepest_descent(forces, x, steps=args.prerelax)
    log(f"[bench] pre-relaxed {len(symbols)} atoms, {args.prerelax} steepest-descent steps in {time.time() - t0:.1f}s")
    if cache:
        np.savez(cache, x=x, prerelax=args.prerelax)
    return symbols, x, cell, types, pbc


def copy_calibration(nbytes, reps=5):
    """A plain device-to-device copy of known size, for scaling HBM counters."""
    import torch

    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "reps": reps}), flush=True)


def cpu_baseline(args, ff):
    """Time the FP64 CPU oracle on a bounded sample of the same workload."""
    import torch

    from oracle import reaxff as orc
    from xreac_amd import workloads

    symbols, x, cell = workloads.water_box(side=args.cpu_side, seed=1)
    t0 = time.time()
    reps = 0
    while True:
        orc.evaluate(ff, symbols, x, cell=cell, pbc=True)
        reps += 1
        if time.time() - t0 > args.cpu_seconds or reps >= 5:
            break
    dt = (time.time() - t0) / reps
    return {
        "value": len(symbols) / dt,
        "unit": "atom-evals/s",
        "cores": int(torch.get_num_threads()),
        "kind": "port",
        "sample": (
            f"{reps} evaluation(s) of a {len(symbols)}-atom water box ({args.cpu_side}^3 waters, same generator and "
            f"density, no pre-relax), E+F+QEq with the FP64 CPU oracle (NumPy/torch-CPU, dense LAPACK QEq); "
            f"{dt:.2f} s per evaluation"
        ),
    }


def build_stream(args):
    """Build the request stream (process pool; call before anything initialises the GPU)."""
    from xreac_amd import workloads

    t0 = time.time()
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    reqs = workloads.request_stream(args.stream_requests, seed=4, workers=workers)
    log(f"[stream] built {len(reqs)} requests in {time.time() - t0:.1f}s with {workers} workers")
    return reqs


def stream_bench(args, reqs):
    """Serving stream: batched FP64, batched mixed and the plain one-request-at-a-time path.

    Requests (seeded, xreac_amd.workloads.stream_plan): atom counts log-uniform
    in [3, 1536]; 60 % water clusters, 30 % CHO mixtures, 10 % charged water
    clusters; 20 % periodic cells; 20 % are 50-step NVE segments (dt 0.25 fs,
    300 K), the rest single evaluations.  An NVE segment counts 51 evaluations.
    """
    import torch

    from xreac_amd.serving import Server

    atoms = np.array([len(r["symbols"]) for r in reqs])
    evals = np.array([r["md_steps"] + 1 for r in reqs])

    def run(server, requests, arrivals=None):
        server.responses = []
        server.atom_evals = 0
        server.steps = 0
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = server.run(requests, arrivals)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        lat = np.array([r.latency for r in out]) * 1e3
        return out, {
            "atom_evals_per_s": server.atom_evals / dt,
            "requests_per_s": len(out) / dt,
            "seconds": dt,
            "requests": len(out),
            "steps": server.steps,
            "errors": int(sum(r.error is not None for r in out)),
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        }

    result = {"requests": len(reqs), "atoms_mean": float(atoms.mean()), "atom_evaluations": int((atoms * evals).sum()),
              "md_fraction": float((evals > 1).mean()), "qeq_tolerance": args.qeq_tol,
              "max_batch_atoms": args.stream_batch_atoms}
    responses = {}
    for mode, precision in (("batched_mixed", "mixed"), ("batched_fp64", "fp64")):
        srv = Server(precision=precision, tol=args.qeq_tol, max_batch_atoms=args.stream_batch_atoms)
        run(srv, reqs[:32])  # warm-up
        out, stats = run(srv, reqs)
        responses[mode] = {r.rid: r for r in out}
        stats["note"] = "all requests queued at t=0 (saturation)"
        result[mode] = stats
        log(f"[stream] {mode}: {stats['atom_evals_per_s']:.4g} atom-evals/s over {stats['seconds']:.1f}s")
    # mixed-precision error against the FP64 path on the single evaluations
    de, df = 0.0, 0.0
    for rid, a in responses["batched_mixed"].items():
        b = responses["batched_fp64"][rid]
        if a.error is None and b.error is None and a.evaluations == 1:
            de = max(de, abs(a.energy - b.energy) / a.atoms)
            df = max(df, float(np.max(np.abs(a.forces - b.forces))))
    result["mixed_vs_fp64"] = {"max_energy_error_per_atom": de, "max_force_error": df,
                               "note": "single evaluations; the FP64 path meets the oracle at 1e-9 rel / 1e-6"}
    # plain: one request at a time, FP64, cold QEq, over a prefix bounded in time
    plain = Server(precision="fp64", warm_start=False, tol=args.qeq_tol, max_systems=1)
    run(plain, reqs[:4])
    plain.responses, plain.atom_evals, plain.steps = [], 0, 0
    torch.cuda.synchronize()
    t = time.perf_counter()
    k = 0
    while k < len(reqs) and time.perf_counter() - t < args.plain_seconds:
        plain.run([reqs[k]])
        k += 1
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    result["plain_fp64"] = {"atom_evals_per_s": plain.atom_evals / dt, "seconds": dt, "requests": k,
                            "note": f"first {k} requests of the same stream, one at a time, cold QEq"}
    result["speedup_batched_mixed_vs_plain"] = result["batched_mixed"]["atom_evals_per_s"] / (plain.atom_evals / dt)
    result["speedup_batched_fp64_vs_plain"] = result["batched_fp64"]["atom_evals_per_s"] / (plain.atom_evals / dt)
    log(f"[stream] plain: {plain.atom_evals / dt:.4g} atom-evals/s over {k} requests")
    # latency under Poisson arrivals at a fraction of the saturated request rate
    rate = args.stream_load * result["batched_mixed"]["requests_per_s"]
    rng = np.random.default_rng(7)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
    srv = Server(precision="mixed", tol=args.qeq_tol, max_batch_atoms=args.stream_batch_atoms)
    run(srv, reqs[:32])
    _, stats = run(srv, reqs, arrivals)
    stats["arrival_rate_per_s"] = rate
    stats["load"] = args.stream_load
    result["poisson_mixed"] = stats
    log(f"[stream] poisson at {rate:.1f} req/s: p50 {stats['latency_ms']['p50']:.1f} ms, "
        f"p99 {stats['latency_ms']['p99']:.1f} ms")
    return result


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc, from profiles/, if recorded."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as fh:
            data = json.load(fh)
        return data.get(name, {}).get("bytes_per_launch")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--side", type=int, default=16, help="waters per box edge (16 -> 12,288 atoms)")
    ap.add_argument("--prerelax", type=int, default=200)
    ap.add_argument("--qeq-tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-side", type=int, default=8, help="CPU sample box edge in waters")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-timers", action="store_true", help="do not bracket kernels with hipEvents")
    ap.add_argument("--no-stream", action="store_true", help="skip the serving-stream measurement")
    ap.add_argument("--only-stream", action="store_true", help="measure only the serving stream (prints its JSON)")
    ap.add_argument("--stream-requests", type=int, default=20000)
    ap.add_argument("--stream-batch-atoms", type=int, default=65536)
    ap.add_argument("--stream-load", type=float, default=0.7, help="Poisson rate as a fraction of saturation")
    ap.add_argument("--plain-seconds", type=float, default=20.0)
    ap.add_argument("--breakdown-steps", type=int, default=2,
                    help="untimed evaluations with every kernel timed, before the timed region")
    ap.add_argument("--geometry-cache", default=None, help="npz file to reuse the pre-relaxed geometry")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(
            "bench.py measures one GPU: this build has no multi-GPU path (requests would shard across GPUs "
            "by whole request, and large systems by spatial domains; not implemented)",
            file=sys.stderr,
        )
        sys.exit(2)

    reqs = None if args.no_stream else build_stream(args)
    if args.only_stream:
        print(json.dumps({"stream": stream_bench(args, reqs)}), flush=True)
        return

    import torch

    from xreac_amd import ForceField
    from xreac_amd.engine import Engine

    ff = ForceField.bundled("ffield.reax.HO.2015")
    eng = Engine(ff, 0)
    symbols, x, cell, types, pbc = build_workload(args, eng, torch)
    n = len(symbols)
    xd = torch.as_tensor(x, device=eng.device).contiguous()

    bufs = eng.outputs(n, properties=True)

    def step():
        return eng.evaluate(xd, types, cell=cell, pbc=pbc, tol=args.qeq_tol, properties=True, out=bufs)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    kernels, dom = {}, None
    if not args.no_timers:
        # Per-kernel breakdown from a short pass before the timed region (every kernel timed).
        eng.profile(True)
        eng.timers(reset=True)
        for _ in range(max(1, args.breakdown_steps)):
            out = step()
        m, nb = out["num_edges"], out["num_bonds"]
        for name, (ms, launches) in eng.timers(reset=True).items():
            if launches == 0:
                continue
            avg_us = 1e3 * ms / launches
            nbd = max(1, args.breakdown_steps)
            row = {"total_ms_per_step": ms / nbd, "launches_per_step": launches / nbd,
                   "avg_us": avg_us}
            b = kernel_bytes(name, n, m, nb)
            if b is not None:
                row["algorithmic_bytes_per_launch"] = b
                row["achieved_GBs"] = b / (avg_us * 1e-6) / 1e9
            kernels[name] = row
        dom = max((k for k in kernels if "achieved_GBs" in kernels[k]), key=lambda k: kernels[k]["total_ms_per_step"])
        # The timed region records events on one in TIMED_EVERY of the dominant kernel's dispatches: events
        # on every dispatch cost about 5 us each, 12 % of a step at 12,288 atoms.
        eng.profile(True, timers=[dom], every=TIMED_EVERY)
        eng.timers(reset=True)
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        out = step()
        lat.append(time.perf_counter() - s0)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    timed = eng.timers(reset=True) if dom else {}
    eng.profile(False)
    m, nb, iters = out["num_edges"], out["num_bonds"], out["qeq_iterations"]
    value = n * args.steps / elapsed

    roofline = None
    if dom:
        ms, launches = timed[dom]
        avg_us = 1e3 * ms / launches
        b = kernel_bytes(dom, n, m, nb)
        achieved = b / (avg_us * 1e-6) / 1e9
        roofline = {
            "kernel": dom,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": read_traffic(dom),
            "avg_launch_us": avg_us,
            "launches_timed": launches,
            "algorithmic_bytes_per_launch": b,
        }

    stream = None
    if not args.no_stream:
        log("[bench] serving stream")
        stream = stream_bench(args, reqs)

    baseline = None
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU oracle on a bounded sample")
        baseline = cpu_baseline(args, ff)

    lat_ms = np.array(lat) * 1e3
    result = {
        "metric": "atom-evals/s (E+F+QEq) batched/mixed vs plain FP64 GPU; p50/p99 latency",
        "value_is": "plain FP64 atom-evals/s on the 12,288-atom water box; the batched/mixed vs plain comparison on "
        "the mixed-size request stream, with p50/p99 latency, is under 'stream'",
        "value": value,
        "unit": "atom-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"periodic water box, {n} atoms ({args.side}^3 H2O, 0.997 g/cm^3, L={cell[0, 0]:.4f} A), "
            f"ffield.reax.HO.2015, {args.prerelax}-step steepest-descent pre-relax",
            "atoms": n,
            "directed_edges": m,
            "active_bonds": nb,
            "qeq_tolerance": args.qeq_tol,
            "qeq_iterations": iters,
            "precision": "plain FP64",
        },
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "roofline": roofline,
        "kernels_breakdown": kernels,
        "stream": stream,
        "cpu_baseline": baseline,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
