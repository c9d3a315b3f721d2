#!/usr/bin/env python
"""Benchmark: ReaxFF energy + forces + QEq evaluations on one MI355X.

Headline (``value``): atom-evaluations/s of the batched mixed-precision form on
the mixed-size serving stream (20,000 seeded requests, 3-1536 atoms, water and
CHO, clusters and periodic cells, single evaluations and 50-step NVE
segments), all requests queued at t = 0.  The mixed form runs FP32 A.m sweeps
with FP64 vectors, dots and true-residual refinement, to the QEq tolerance that
keeps every single evaluation within 1e-4 kcal/mol/atom (energy) and 1e-3
kcal/mol/A (forces) of the FP64 path at tolerance 1e-10; the measured
errors are printed beside it.  An atom-evaluation is one atom through E + F +
QEq; an NVE segment counts its 51 evaluations.

Beside it, in the same JSON line: the same stream batched in FP64 (tolerance
1e-10, and the mixed form's tolerance), the plain FP64 path (one request at a
time, cold QEq), Poisson arrivals at three loads with p50/p99 latency, the
12,288-atom water box (plain FP64, tolerance 1e-10, the roofline of its
dominant kernel and its error against the committed oracle fixture), the
3,150-atom CHO mixture and the 24,360-atom anatase/water slab (mixed against
FP64 at tolerance 1e-10), NVE steps of the 192-atom cluster run unrecorded and
as recorded HIP-graph replays, and the FP64 CPU oracle timed on this host.
Positions are resident in HBM when every timed region starts.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TIMED_EVERY = 8  # timed region: events on one dispatch in 8 of the dominant kernel
MIXED_TOL = 5e-8  # QEq tolerance of the mixed form (errors measured against FP64 at 1e-10 below)
FIXTURE = os.path.join(ROOT, "tests", "golden", "water_box_12288.npz")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_bytes(name, n, m, nb):
    """Algorithmic HBM bytes per launch of a timed kernel class.

    cg_iteration: one CG iteration over both right-hand sides: the QEq matrix
    read once in half storage (8-B value + 4-B column per unordered pair,
    P = M/2) and 152 B per atom of (s, t) vector traffic (p, A p, x, r and the
    diagonal), i.e. 12 P + 152 N.
    """
    P = m / 2
    if name == "cg_iteration":
        return 12 * P + 152 * n
    if name == "nonbonded":
        return 16 * P + 56 * n  # half list (j + image code) + positions, types, charges, forces
    if name == "edge":
        return 8 * P + 12 * P + 40 * n  # half list read, H write, positions
    if name == "neighbor_fill":
        return 16 * P + 32 * n
    return None


def implementation_bytes(name, n, m):
    """Bytes this implementation's k_cg_iteration touches per launch (not the roofline's):
    H stored for both edge directions (8 + 4 B per directed edge) and 436 B per atom
    (eight recurrence vectors read and written, m in and out, the block-Jacobi row,
    eta, block placement and row pointer)."""
    return 12 * m + 436 * n if name == "cg_iteration" else None


# ------------------------------------------------------------------ CPU oracle
def cpu_child(spec):
    """Runs in a child process (thread limits set in its environment before numpy loads)."""
    import torch

    from oracle import reaxff as orc
    from xreac_amd import ForceField, workloads

    torch.set_num_threads(spec["threads"])
    ff = ForceField.bundled("ffield.reax.HO.2015")
    if spec["system"] == "water_box_12288":
        s, x, cell = workloads.prerelaxed_water_box()
    else:
        s, x, cell = workloads.water_box(side=spec["side"], seed=1)
    times = []
    t_end = time.time() + spec["seconds"]
    for k in range(spec["calls"] + spec["warmup"]):
        t0 = time.perf_counter()
        orc.evaluate(ff, s, x, cell=cell, pbc=True)
        if k >= spec["warmup"]:
            times.append(time.perf_counter() - t0)
        if time.time() > t_end and times:
            break
    print(json.dumps({"atoms": len(s), "seconds": times}), flush=True)


def run_cpu(spec):
    env = dict(os.environ)
    threads = str(spec["threads"])
    # the reference's convention for thread-count control (scripts/benchmark_lammps.py:19-31)
    for var in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "VECLIB_MAXIMUM_THREADS",
                "NUMEXPR_NUM_THREADS"):
        env[var] = threads
    env["HIP_VISIBLE_DEVICES"] = ""  # the oracle never touches the GPU
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", json.dumps(spec)], env=env,
                         capture_output=True, text=True, timeout=1200)
    if out.returncode != 0:
        raise RuntimeError(f"CPU oracle child failed: {out.stderr[-2000:]}")
    return json.loads(out.stdout.strip().splitlines()[-1])


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args):
    """The FP64 CPU oracle (reference-structured: NumPy per-edge math, dense LAPACK QEq, autograd forces)
    on this host: the 12,288-atom box at all cores, and a 1,536-atom box at one thread."""
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    t0 = time.time()
    full = run_cpu(dict(system="water_box_12288", threads=cores, calls=args.cpu_calls, warmup=0, seconds=0))
    log(f"[cpu] 12,288-atom box, {cores} threads: {np.median(full['seconds']):.1f} s per evaluation")
    one = run_cpu(dict(system="water_box", side=8, threads=1, calls=3, warmup=1, seconds=args.cpu_seconds))
    log(f"[cpu] 1,536-atom box, 1 thread: {np.median(one['seconds']):.2f} s per evaluation "
        f"({time.time() - t0:.0f} s in all)")
    t_full, t_one = float(np.median(full["seconds"])), float(np.median(one["seconds"]))
    return {
        "value": full["atoms"] / t_full,
        "unit": "atom-evals/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"{len(full['seconds'])} evaluation(s) of the pre-relaxed 12,288-atom water box (the workload of "
                   f"'water_box_12288'), E+F+QEq with the FP64 CPU oracle (NumPy/torch-CPU per-edge math, dense "
                   f"LAPACK QEq, autograd forces), {cores} threads, median {t_full:.1f} s per evaluation"),
        "cpu_model": cpu_model(),
        "single_thread": {"value": one["atoms"] / t_one, "unit": "atom-evals/s", "cores": 1,
                          "sample": f"1,536-atom water box (8^3 waters), 1 warm-up + median of "
                                    f"{len(one['seconds'])} calls: {t_one:.2f} s per evaluation"},
    }


# ------------------------------------------------------------------ serving stream
def build_stream(args):
    """Build the request stream (process pool; call before anything initialises the GPU)."""
    from xreac_amd import workloads

    t0 = time.time()
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    reqs = workloads.request_stream(args.stream_requests, seed=4, workers=workers)
    log(f"[stream] built {len(reqs)} requests in {time.time() - t0:.1f}s with {workers} workers")
    return reqs


def stream_bench(args, reqs):
    """The serving stream: batched mixed (the headline), batched FP64, plain FP64 and Poisson loads.

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
            "ms_per_step": 1e3 * dt / max(1, server.steps),
            "errors": int(sum(r.error is not None for r in out)),
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        }

    result = {"requests": len(reqs), "atoms_mean": float(atoms.mean()), "atom_evaluations": int((atoms * evals).sum()),
              "md_fraction": float((evals > 1).mean()), "max_batch_atoms": args.stream_batch_atoms}
    responses = {}
    modes = (("batched_mixed", "mixed", args.mixed_tol), ("batched_fp64", "fp64", args.qeq_tol),
             ("batched_fp64_at_mixed_tol", "fp64", args.mixed_tol))
    for mode, precision, tol in modes:
        srv = Server(precision=precision, tol=tol, max_batch_atoms=args.stream_batch_atoms,
                     recorded_steps=args.stream_recorded)
        # warm-up on the first 3,000 requests: enough for full-size batches, so the engines' buffers have
        # grown to their largest before the timed run (else the first mode timed pays the growth)
        run(srv, reqs[:3000])
        out, stats = run(srv, reqs)
        responses[mode] = {r.rid: r for r in out}
        stats.update(precision=precision, qeq_tolerance=tol, note="all requests queued at t=0 (saturation); "
                     "latency here is mostly queueing")
        result[mode] = stats
        log(f"[stream] {mode}: {stats['atom_evals_per_s']:.4g} atom-evals/s over {stats['seconds']:.1f}s")
    # every single evaluation of the mixed form (and of FP64 at the mixed tolerance) against FP64 at 1e-10
    for mode in ("batched_mixed", "batched_fp64_at_mixed_tol"):
        de = df = dq = 0.0
        for rid, a in responses[mode].items():
            b = responses["batched_fp64"][rid]
            if a.error is None and b.error is None and a.evaluations == 1:
                de = max(de, abs(a.energy - b.energy) / a.atoms)
                df = max(df, float(np.max(np.abs(a.forces - b.forces))))
                dq = max(dq, float(np.max(np.abs(a.charges - b.charges))))
        result[mode]["error_vs_fp64"] = {
            "max_energy_error_per_atom": de, "max_force_error": df, "max_charge_error": dq,
            "targets": {"energy_per_atom": 1e-4, "force": 1e-3},
            "note": "every single evaluation of the stream against batched FP64 at tolerance 1e-10, which meets "
                    "the FP64 oracle at 1e-9 relative energy, 1e-6 forces, 1e-7 charges (pytest -m gpu)",
        }
    # plain: one request at a time, FP64, cold QEq, over a prefix bounded in time
    plain = Server(precision="fp64", warm_start=False, tol=args.qeq_tol, max_systems=1, recorded_steps=False)
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
    plain_rate = plain.atom_evals / dt
    result["plain_fp64"] = {"atom_evals_per_s": plain_rate, "seconds": dt, "requests": k,
                            "qeq_tolerance": args.qeq_tol,
                            "note": f"first {k} requests of the same stream, one at a time, cold QEq"}
    for mode in ("batched_mixed", "batched_fp64", "batched_fp64_at_mixed_tol"):
        result[mode]["vs_plain_fp64"] = result[mode]["atom_evals_per_s"] / plain_rate
    log(f"[stream] plain: {plain_rate:.4g} atom-evals/s over {k} requests")
    # latency under Poisson arrivals at three fractions of the saturated request rate (mixed form)
    sub = reqs[: args.poisson_requests]
    result["poisson_mixed"] = []
    for load in args.stream_loads:
        rate = load * result["batched_mixed"]["requests_per_s"]
        rng = np.random.default_rng(7)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, len(sub)))
        srv = Server(precision="mixed", tol=args.mixed_tol, max_batch_atoms=args.stream_batch_atoms,
                     recorded_steps=args.stream_recorded)
        run(srv, reqs[:32])
        _, stats = run(srv, sub, arrivals)
        stats.update(arrival_rate_per_s=rate, load=load, qeq_tolerance=args.mixed_tol)
        result["poisson_mixed"].append(stats)
        log(f"[stream] poisson load {load} ({rate:.0f} req/s): p50 {stats['latency_ms']['p50']:.1f} ms, "
            f"p99 {stats['latency_ms']['p99']:.1f} ms")
    return result


# ------------------------------------------------------------------ single systems
def timed_evaluations(eng, x, types, cell, steps, warmup, tol, precision, properties=True, timer=None,
                      neighbor_skin=0.0):
    """Mean ms per evaluation over `steps` evaluations (after `warmup`), and the last result."""
    import torch

    xd = torch.as_tensor(x, device=eng.device).contiguous()
    bufs = eng.outputs(len(x), properties=properties)

    def step():
        return eng.evaluate(xd, types, cell=cell, pbc=(True,) * 3, tol=tol, precision=precision,
                            properties=properties, out=bufs, neighbor_skin=neighbor_skin)

    for _ in range(warmup):
        step()
    if timer:
        eng.profile(True, timers=[timer], every=TIMED_EVERY)
        eng.timers(reset=True)
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        out = step()
        lat.append(time.perf_counter() - s0)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    timed = eng.timers(reset=True) if timer else None
    if timer:
        eng.profile(False)
    return 1e3 * elapsed / steps, np.array(lat) * 1e3, out, timed


def breakdown(eng, x, types, cell, tol, n_steps=2):
    """Per-kernel times from untimed evaluations with every kernel timed."""
    import torch

    xd = torch.as_tensor(x, device=eng.device).contiguous()
    eng.profile(True)
    eng.timers(reset=True)
    for _ in range(max(1, n_steps)):
        out = eng.evaluate(xd, types, cell=cell, pbc=(True,) * 3, tol=tol, properties=True)
    rows = {}
    m, nb, n = out["num_edges"], out["num_bonds"], len(x)
    for name, (ms, launches) in eng.timers(reset=True).items():
        if launches == 0:
            continue
        avg_us = 1e3 * ms / launches
        row = {"total_ms_per_step": ms / n_steps, "launches_per_step": launches / n_steps, "avg_us": avg_us}
        b = kernel_bytes(name, n, m, nb)
        if b is not None:
            row["algorithmic_bytes_per_launch"] = b
            row["achieved_GBs"] = b / (avg_us * 1e-6) / 1e9
        rows[name] = row
    eng.profile(False)
    return rows


def errors(a, b, n):
    """Energy (per atom and relative), force and charge differences of two evaluations."""
    ea, eb = float(a["components"].sum()), float(b["components"].sum())
    return {
        "energy_per_atom": abs(ea - eb) / n,
        "energy_relative": abs(ea - eb) / abs(eb),
        "max_force": float(np.max(np.abs(a["forces"].cpu().numpy() - b["forces"].cpu().numpy()))),
        "max_charge": float(np.max(np.abs(a["charges"].cpu().numpy() - b["charges"].cpu().numpy()))),
    }


def water_box_bench(args):
    """The pre-relaxed 12,288-atom box, plain FP64 at tolerance 1e-10: timed steps, roofline, fixture errors."""
    import hashlib

    import torch

    from xreac_amd import ForceField, workloads
    from xreac_amd.engine import Engine

    ff = ForceField.bundled("ffield.reax.HO.2015")
    eng = Engine(ff, 0)
    s, x, cell = workloads.prerelaxed_water_box()
    n = len(s)
    types = torch.as_tensor(eng.type_ids(s), device=eng.device)
    kernels = breakdown(eng, x, types, cell, args.qeq_tol, args.breakdown_steps) if not args.no_timers else {}
    dom = max((k for k in kernels if "achieved_GBs" in kernels[k]), key=lambda k: kernels[k]["total_ms_per_step"],
              default=None)
    ms, lat, out, timed = timed_evaluations(eng, x, types, cell, args.steps, args.warmup, args.qeq_tol, "fp64",
                                            timer=dom)
    m, nb, iters = out["num_edges"], out["num_bonds"], out["qeq_iterations"]
    roofline = None
    if dom:
        t_ms, launches = timed[dom]
        avg_us = 1e3 * t_ms / launches
        b = kernel_bytes(dom, n, m, nb)
        achieved = b / (avg_us * 1e-6) / 1e9
        roofline = {
            "kernel": dom,
            "workload": "water_box_12288, plain FP64",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": read_traffic(dom),
            "avg_launch_us": avg_us,
            "launches_timed": launches,
            "algorithmic_bytes_per_launch": b,
            "implementation_bytes_per_launch": implementation_bytes(dom, n, m),
        }
    result = {
        "workload": f"periodic water box, {n} atoms (16^3 H2O, 0.997 g/cm^3, L={cell[0, 0]:.4f} A), "
                    f"ffield.reax.HO.2015, 200-step steepest-descent pre-relax (data/water_box_12288_prerelaxed.npz)",
        "atoms": n, "directed_edges": m, "active_bonds": nb, "qeq_tolerance": args.qeq_tol, "qeq_iterations": iters,
        "precision": "plain FP64", "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "atom_evals_per_s": n / (ms * 1e-3),
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "kernels_breakdown": kernels,
    }
    # error of this evaluation against the committed FP64 CPU oracle results for the same coordinates
    if os.path.exists(FIXTURE):
        ref = np.load(FIXTURE)
        sub = ref["subset"]
        comps = out["components"]
        e = float(comps.sum())
        row_ptr, j, S = eng.neighbors(n, m)
        i = np.repeat(np.arange(n), np.diff(row_ptr))
        h = hashlib.sha256()
        for a in (i, j, S):
            h.update(np.ascontiguousarray(a, dtype=np.int64).tobytes())
        result["error_vs_oracle"] = {
            "energy_relative": abs(e - float(ref["energy"])) / abs(float(ref["energy"])),
            "max_component": float(np.max(np.abs(comps - ref["components"]))),
            "max_force_subset": float(np.max(np.abs(out["forces"].cpu().numpy()[sub] - ref["forces"]))),
            "max_charge_subset": float(np.max(np.abs(out["charges"].cpu().numpy()[sub] - ref["charges"]))),
            "neighbor_list_identical": h.hexdigest() == str(ref["neighbor_sha256"]),
            "bond_count_identical": nb == int(ref["num_bonds"]),
            "note": f"against tests/golden/water_box_12288.npz (oracle, same coordinates); forces and charges on "
                    f"its {len(sub)} seeded atoms",
        }
    # mixed precision on the same box: at 1e-10 and at the mixed form's tolerance
    base = out
    fp64_same = timed_evaluations(eng, x, types, cell, max(3, args.steps // 2), 1, args.mixed_tol, "fp64")
    for label, tol in (("mixed_tol_1e-10", args.qeq_tol), ("mixed_at_mixed_tol", args.mixed_tol)):
        ms_m, _, out_m, _ = timed_evaluations(eng, x, types, cell, max(3, args.steps // 2), 1, tol, "mixed")
        result[label] = {"ms_per_step": ms_m, "qeq_tolerance": tol, "qeq_iterations": out_m["qeq_iterations"],
                         "error_vs_fp64": errors(out_m, base, n)}
    result["fp64_at_mixed_tol"] = {"ms_per_step": fp64_same[0], "qeq_tolerance": args.mixed_tol,
                                   "qeq_iterations": fp64_same[2]["qeq_iterations"],
                                   "error_vs_fp64": errors(fp64_same[2], base, n)}
    # a kept skin list (the reference's ASE neighbor_skin, 0.3 A): after its first build every call filters it
    builds0 = eng.neighbor_list_builds
    ms_s, _, out_s, _ = timed_evaluations(eng, x, types, cell, args.steps, 1, args.qeq_tol, "fp64",
                                          neighbor_skin=0.3)
    same = all(np.array_equal(out_s[k].cpu().numpy(), base[k].cpu().numpy()) for k in ("forces", "charges"))
    result["fp64_kept_skin_list"] = {"ms_per_step": ms_s, "neighbor_skin": 0.3,
                                     "list_builds": eng.neighbor_list_builds - builds0,
                                     "evaluations": args.steps + 1, "bitwise_equal_to_fresh_list": bool(same),
                                     "note": "same positions every call (the reuse path only); an NVE segment "
                                             "rebuilds when an atom has moved more than the skin"}
    eng.close()
    return result, roofline


def system_bench(args, name, ffname, build, steps):
    """FP64 at 1e-10 against the mixed form on one system (the CHO mixture, the anatase/water slab)."""
    import torch

    from xreac_amd import ForceField
    from xreac_amd.engine import Engine

    eng = Engine(ForceField.bundled(ffname), 0)
    s, x, cell = build()
    n = len(s)
    types = torch.as_tensor(eng.type_ids(s), device=eng.device)
    ms, _, base, _ = timed_evaluations(eng, x, types, cell, steps, 2, args.qeq_tol, "fp64")
    kernels = breakdown(eng, x, types, cell, args.qeq_tol, 1) if not args.no_timers else {}
    result = {"workload": name, "atoms": n, "directed_edges": base["num_edges"], "active_bonds": base["num_bonds"],
              "fp64": {"ms_per_step": ms, "atom_evals_per_s": n / (ms * 1e-3), "qeq_tolerance": args.qeq_tol,
                       "qeq_iterations": base["qeq_iterations"]},
              "cg_iteration_avg_us": kernels.get("cg_iteration", {}).get("avg_us")}
    for label, tol in (("mixed_tol_1e-10", args.qeq_tol), ("mixed_at_mixed_tol", args.mixed_tol)):
        ms_m, _, out_m, _ = timed_evaluations(eng, x, types, cell, steps, 2, tol, "mixed")
        result[label] = {"ms_per_step": ms_m, "atom_evals_per_s": n / (ms_m * 1e-3), "qeq_tolerance": tol,
                         "qeq_iterations": out_m["qeq_iterations"], "error_vs_fp64": errors(out_m, base, n)}
    eng.close()
    log(f"[{name}] fp64 {ms:.3f} ms, mixed {result['mixed_tol_1e-10']['ms_per_step']:.3f} ms (tol 1e-10)")
    return result


def recorded_bench(args, steps=100, warmup=5):
    """NVE steps run unrecorded (host-issued launches, QEq read-backs) and as recorded graph replays.

    Each step is velocity-Verlet stage 0, extrapolated QEq guesses, the batch
    evaluation (FP64, warm start, tolerance 1e-10) and stage 1, as the server
    runs an NVE segment; the two runs start from the same state and must end
    bitwise equal.  Cases: the 64-water, 192-atom cluster, alone and as 64
    seeded copies in one batch.
    """
    import torch

    from xreac_amd import ForceField, workloads
    from xreac_amd.engine import Engine

    ff = ForceField.bundled("ffield.reax.HO.2015")
    eng = Engine(ff, 0, stream=torch.cuda.Stream(0))
    out = {}
    for label, copies in (("water_cluster_192", 1), ("water_cluster_192_x64", 64)):
        systems = [workloads.water_cluster(seed=k) for k in range(copies)]
        sym = [a for s, _ in systems for a in s]
        x0 = np.vstack([x for _, x in systems])
        offsets = np.array([0] + list(np.cumsum([len(s) for s, _ in systems])), dtype=np.int32)
        masses = np.array([ff.atoms[a]["mass"] for a in sym])
        v0 = workloads.maxwell_boltzmann(masses, 300.0, seed=3)
        n, ns = len(sym), copies
        finals, rates = {}, {}
        with torch.cuda.stream(eng.stream):
            dev = eng.device
            f64 = dict(dtype=torch.float64, device=dev)
            for mode in ("unrecorded", "recorded"):
                x = torch.as_tensor(x0, **f64).contiguous()
                v = torch.as_tensor(v0, **f64).contiguous()
                t = torch.as_tensor(eng.type_ids(sym), device=dev)
                o = eng.batch_outputs(n, qeq_state=True)
                o["forces"].zero_()
                qh = torch.zeros((3, n, 2), **f64)
                kin = torch.zeros(ns, **f64)
                off = torch.as_tensor(offsets, device=dev)
                dt = torch.zeros(ns, **f64)
                order = torch.zeros(ns, dtype=torch.int32, device=dev)
                # the first evaluation of the segment (nothing moves), then the steps
                eng.evaluate_batch(x, t, offsets, tol=1e-10, warm_start=True, out=o)
                dt.fill_(0.25)
                step = None
                if mode == "recorded":
                    step = eng.record_step(x, t, offsets, out=o, tol=1e-10, warm_start=True, dt_fs=dt, velocities=v,
                                           kinetic=kin, qeq_order=order, qeq_history=qh)
                torch.cuda.synchronize()
                time.sleep(0.2)  # a gap that separates the modes in a kernel trace
                t0 = 0.0
                for k in range(warmup + steps):
                    if k == warmup:
                        torch.cuda.synchronize()
                        t0 = time.perf_counter()
                    order.fill_(min(k + 1, 4) - 1)
                    if step is not None:
                        step.run()
                    else:
                        eng.verlet(off, t, dt, x, v, o["forces"], 0)
                        eng.qeq_extrapolate(off, order, o["qeq_state"], qh)
                        eng.evaluate_batch(x, t, offsets, tol=1e-10, warm_start=True, out=o)
                        eng.verlet(off, t, dt, x, v, o["forces"], 1, kin)
                    kin.cpu()  # the server reads the kinetic energies back every step
                torch.cuda.synchronize()
                sec = time.perf_counter() - t0
                rates[mode] = 1e3 * sec / steps
                finals[mode] = (x.cpu().numpy(), o["forces"].cpu().numpy())
                if step is not None:
                    replays, fallbacks, recordings = step.stats()
                    step.close()
        same = all(np.array_equal(a, b) for a, b in zip(finals["unrecorded"], finals["recorded"]))
        out[label] = {"atoms": n, "systems": ns, "steps": steps,
                      "ms_per_step_unrecorded": rates["unrecorded"], "ms_per_step_recorded": rates["recorded"],
                      "speedup": rates["unrecorded"] / rates["recorded"], "bitwise_equal": same,
                      "replays": replays, "fallbacks": fallbacks, "recordings": recordings}
        log(f"[recorded] {label}: {rates['unrecorded']:.3f} -> {rates['recorded']:.3f} ms/step "
            f"(bitwise equal: {same}; replays {replays}, fallbacks {fallbacks})")
    eng.close()
    return out


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


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed evaluations of the 12,288-atom box")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--qeq-tol", type=float, default=1e-10)
    ap.add_argument("--mixed-tol", type=float, default=MIXED_TOL)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-calls", type=int, default=1, help="oracle evaluations of the 12,288-atom box")
    ap.add_argument("--cpu-seconds", type=float, default=30.0, help="bound of the single-thread sample")
    ap.add_argument("--no-timers", action="store_true", help="do not bracket kernels with hipEvents")
    ap.add_argument("--no-stream", action="store_true", help="skip the serving stream (value is then the box's)")
    ap.add_argument("--no-systems", action="store_true", help="skip the CHO mixture and the anatase slab")
    ap.add_argument("--no-recorded", action="store_true", help="skip the recorded-step (HIP graph) NVE comparison")
    ap.add_argument("--only-recorded", action="store_true", help="measure only the recorded-step NVE comparison")
    ap.add_argument("--only-box", action="store_true", help="measure only the 12,288-atom box")
    ap.add_argument("--only-stream", action="store_true", help="measure only the serving stream (for profiling)")
    ap.add_argument("--stream-requests", type=int, default=20000)
    ap.add_argument("--stream-batch-atoms", type=int, default=262144)
    ap.add_argument("--stream-loads", type=float, nargs="+", default=[0.5, 0.7, 0.9],
                    help="Poisson rates as fractions of the saturated request rate")
    ap.add_argument("--poisson-requests", type=int, default=5000)
    ap.add_argument("--stream-recorded", action="store_true",
                    help="serve the stream with recorded steps (HIP-graph replays of unchanged groups)")
    ap.add_argument("--plain-seconds", type=float, default=15.0)
    ap.add_argument("--breakdown-steps", type=int, default=2,
                    help="untimed evaluations with every kernel timed, before the timed region")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (HBM counter calibration) and exit")
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_child:
        cpu_child(json.loads(args.cpu_child))
        return
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
    if args.only_box:
        args.no_stream = args.no_systems = args.no_cpu_baseline = args.no_recorded = True
    if args.only_stream:
        args.no_systems = args.no_cpu_baseline = args.no_recorded = True

    if args.only_recorded:
        print(json.dumps({"recorded_steps": recorded_bench(args)}), flush=True)
        return
    # Process pools and CPU children start before this process touches the GPU.
    reqs = None if args.no_stream else build_stream(args)
    baseline = None
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU oracle")
        baseline = cpu_baseline(args)

    box = roofline = None
    if not args.only_stream:
        box, roofline = water_box_bench(args)
        log(f"[box] {box['ms_per_step']:.3f} ms per evaluation, {box['atom_evals_per_s']:.4g} atom-evals/s")
    systems = {}
    if not args.no_systems:
        from xreac_amd import workloads

        systems["cho_mixture_3150"] = system_bench(
            args, "cho_mixture_3150: 300 CH4 + 300 O2 + 150 H2O + 100 CH3OH, 40 A periodic cube, ffield.reax.CHO.2008",
            "ffield.reax.CHO.2008", workloads.cho_mixture, 10)
        systems["anatase_slab_24360"] = system_bench(
            args, "anatase_slab_24360: anatase 16x16x5 + 3000 H2O + 15 A vacuum, periodic, ffield.reax.TiOH.Kim2013",
            "ffield.reax.TiOH.Kim2013", workloads.anatase_water_slab, 5)
    recorded = None if args.no_recorded else recorded_bench(args)
    stream = None if args.no_stream else stream_bench(args, reqs)

    if stream is not None:
        head = stream["batched_mixed"]
        value, unit_steps, ms_step = head["atom_evals_per_s"], head["steps"], head["ms_per_step"]
        workload = (f"serving stream: {stream['requests']} seeded requests, 3-1536 atoms (log-uniform, mean "
                    f"{stream['atoms_mean']:.0f}), 60 % water / 30 % CHO / 10 % charged water clusters, 20 % "
                    f"periodic, 20 % 50-step NVE; all queued at t=0, batched mixed precision at QEq tolerance "
                    f"{args.mixed_tol:g}")
        value_is = ("atom-evals/s of the batched mixed-precision form (FP32 A.m sweeps, FP64 vectors, dots and "
                    "refinement) over the whole stream; 'steps' are server steps (one batched evaluation of every "
                    "active request each); the FP64 and plain forms, Poisson latency, the single systems and the "
                    "CPU oracle are under 'stream', 'water_box_12288', 'systems' and 'cpu_baseline'")
        dtype = "f64 (f32 QEq sweeps)"
    else:
        value, unit_steps, ms_step = box["atom_evals_per_s"], box["steps"], box["ms_per_step"]
        workload = box["workload"]
        value_is = "atom-evals/s of the 12,288-atom box, plain FP64 (the stream was skipped)"
        dtype = "f64"
    result = {
        "metric": "atom-evals/s (E+F+QEq) batched/mixed vs plain FP64 GPU; p50/p99 latency",
        "value_is": value_is,
        "value": value,
        "unit": "atom-evals/s",
        "n_gpus": 1,
        "steps": unit_steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded generators; force fields bundled in data/)",
        "config": {"workload": workload, "qeq_tolerance_mixed": args.mixed_tol, "qeq_tolerance_fp64": args.qeq_tol,
                   "max_batch_atoms": args.stream_batch_atoms},
        "speedup_batched_mixed_vs_plain_fp64": stream["batched_mixed"]["vs_plain_fp64"] if stream else None,
        "latency_ms": stream["poisson_mixed"][1]["latency_ms"] if stream and len(stream["poisson_mixed"]) > 1
        else None,
        "roofline": roofline,
        "stream": stream,
        "water_box_12288": box,
        "systems": systems,
        "recorded_steps": recorded,
        "cpu_baseline": baseline,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
