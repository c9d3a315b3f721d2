"""Continuous-batching server on an MI355X.

Every request of a mixed stream (water and CHO, clusters and periodic cells,
charged requests, single evaluations and short NVE segments) must give the
same result served in the mix as served alone, and single evaluations must
meet the FP64 oracle bars (plain FP64 QEq) or the mixed-precision targets
(1e-4 kcal/mol/atom energy, 1e-3 kcal/mol/A forces).
"""

import numpy as np
import pytest

from oracle import dynamics as od
from oracle import reaxff as orc
from xreac_amd import ForceField, workloads

pytestmark = pytest.mark.gpu


def _requests(count, seed, md_steps=4, max_atoms=160, relax_every=0):
    plan = workloads.stream_plan(count, seed=seed, max_atoms=max_atoms, md_fraction=0.3, md_steps=md_steps)
    reqs = [workloads.build_request(p) for p in plan]
    if relax_every:
        for r in reqs[::relax_every]:
            if r["md_steps"] == 0:
                r["relax"] = True
    return reqs


def _serve(requests, **kw):
    from xreac_amd.serving import Server

    srv = Server(**kw)
    out = srv.run(requests)
    return {r.rid: r for r in out}, srv


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_mix_equals_alone(precision):
    reqs = _requests(24, seed=21, relax_every=5)
    assert any(r.get("relax") for r in reqs)
    print(f"[serving] {len(reqs)} requests, {precision}", flush=True)
    # a small atom cap forces admission over several steps while others run
    kw = dict(precision=precision, relax_iterations=8)
    mixed, srv = _serve(reqs, max_batch_atoms=600, **kw)
    assert srv.steps >= 6, srv.steps
    assert len(mixed) == len(reqs)
    for req in reqs:
        alone, _ = _serve([req], **kw)
        a, m = alone[req["rid"]], mixed[req["rid"]]
        assert m.error is None and a.error is None, (m.error, a.error)
        assert m.evaluations == a.evaluations
        if req.get("relax"):
            assert 1 <= m.evaluations <= 9
            np.testing.assert_array_equal(m.positions, a.positions)
            assert m.converged == a.converged
        else:
            assert m.evaluations == req["md_steps"] + 1
        np.testing.assert_array_equal(m.forces, a.forces)
        np.testing.assert_array_equal(m.charges, a.charges)
        np.testing.assert_array_equal(m.components, a.components)
        if req["md_steps"]:
            np.testing.assert_array_equal(m.positions, a.positions)
            np.testing.assert_array_equal(m.velocities, a.velocities)
            np.testing.assert_array_equal(m.potential, a.potential)


def test_admission_respects_atom_cap():
    """All requests queued at once: every evaluated batch stays within max_batch_atoms (a request larger
    than the cap is admitted alone)."""
    from xreac_amd.serving import Server

    reqs = _requests(40, seed=17, max_atoms=160)
    cap = 400
    srv = Server(max_batch_atoms=cap)
    sizes = []
    for eng in srv.engines.values():
        run = eng.evaluate_batch

        def timed(*a, _run=run, **k):
            sizes.append(int(a[2][-1]))
            return _run(*a, **k)

        eng.evaluate_batch = timed
    out = srv.run(reqs)
    assert len(out) == len(reqs)
    biggest = max(len(r["symbols"]) for r in reqs)
    print(f"[serving] {len(sizes)} batches, largest {max(sizes)} atoms (cap {cap})", flush=True)
    assert max(sizes) <= max(cap, biggest)
    assert len(sizes) > len(srv.engines)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_served_evaluations_match_oracle(precision):
    reqs = [r for r in _requests(40, seed=5, max_atoms=90) if r["md_steps"] == 0][:8]
    res, _ = _serve(reqs, precision=precision, tol=1e-10)
    ffs = {}
    for req in reqs:
        ff = ffs.setdefault(req["force_field"], ForceField.bundled(req["force_field"]))
        ref = orc.evaluate(ff, req["symbols"], req["positions"], cell=req["cell"], pbc=req["pbc"],
                           total_charge=req["total_charge"])
        r = res[req["rid"]]
        n = len(req["symbols"])
        print(f"[serving] oracle check rid={req['rid']} n={n}", flush=True)
        if precision == "fp64":
            assert abs(r.energy - ref["energy"]) <= 1e-9 * max(1.0, abs(ref["energy"]))
            np.testing.assert_allclose(r.forces, ref["forces"], rtol=0, atol=1e-6)
            np.testing.assert_allclose(r.charges, ref["charges"], rtol=0, atol=1e-7)
        else:
            assert abs(r.energy - ref["energy"]) <= 1e-4 * n
            np.testing.assert_allclose(r.forces, ref["forces"], rtol=0, atol=1e-3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_served_large_requests_match_oracle(precision):
    """Requests near the stream's 1,536-atom cap, served in one mix with small ones: a periodic water cell
    (1,527 atoms), a periodic CHO cell (1,393), a charged water cluster (1,058) and a CHO cluster (1,231)
    of the seeded stream, against the FP64 oracle."""
    plan = workloads.stream_plan(400, seed=4)
    big = [workloads.build_request(plan[rid]) for rid in (230, 376, 48, 272)]
    assert [len(r["symbols"]) for r in big] == [1527, 1393, 1058, 1231]
    assert big[0]["cell"] is not None and big[1]["cell"] is not None and big[2]["total_charge"] == -1.0
    small = [r for r in _requests(30, seed=6, max_atoms=60) if r["md_steps"] == 0][:6]
    for k, r in enumerate(small):
        r["rid"] = 10**6 + k
    res, _ = _serve(big + small, precision=precision, tol=1e-10)
    ffs = {}
    for req in big:
        ff = ffs.setdefault(req["force_field"], ForceField.bundled(req["force_field"]))
        n = len(req["symbols"])
        print(f"[serving] oracle check rid={req['rid']} n={n} {req['force_field']} {precision}", flush=True)
        ref = orc.evaluate(ff, req["symbols"], req["positions"], cell=req["cell"], pbc=req["pbc"],
                           total_charge=req["total_charge"])
        r = res[req["rid"]]
        assert r.error is None
        if precision == "fp64":
            assert abs(r.energy - ref["energy"]) <= 1e-9 * max(1.0, abs(ref["energy"]))
            np.testing.assert_allclose(r.forces, ref["forces"], rtol=0, atol=1e-6)
            np.testing.assert_allclose(r.charges, ref["charges"], rtol=0, atol=1e-7)
        else:
            assert abs(r.energy - ref["energy"]) <= 1e-4 * n
            np.testing.assert_allclose(r.forces, ref["forces"], rtol=0, atol=1e-3)


def test_nve_segment_stays_bounded():
    """50-step NVE segments: fixed-charge forces are not the exact gradient of the QEq energy (14.4 in H
    against 332.06371/23.02 in the energy), so the total drifts a little; it must stay a small fraction of
    the kinetic energy."""
    reqs = [r for r in _requests(30, seed=9, md_steps=50, max_atoms=120) if r["md_steps"]][:3]
    res, _ = _serve(reqs)
    for req in reqs:
        r = res[req["rid"]]
        total = r.potential + r.kinetic
        print(f"[serving] NVE rid={req['rid']} n={r.atoms} range={np.ptp(total):.3e} "
              f"mean KE={r.kinetic.mean():.3e}", flush=True)
        assert r.evaluations == 51
        assert np.all(np.isfinite(total))
        assert np.ptp(total) < 0.02 * r.kinetic.mean() + 0.1


def test_nve_conserves_energy_with_exact_gradient():
    """With charge-response forces (the exact gradient of the QEq-equilibrated energy) velocity Verlet
    conserves the total energy to O(dt^2): the spread of the total over 50 steps of 0.25 fs stays below
    1e-2 of the kinetic energy (O-H vibrations, period about 9 fs, set the scale; 2.6e-3 for the 192-atom
    cluster), and halving the step over the same time divides it by four.  (Fixed-charge forces are not a gradient; their
    spread does not shrink that way.)"""
    import copy

    reqs = []
    for rid, (s, x, cell) in enumerate((
        (*workloads.water_cluster(), None),  # the 64-water, 192-atom cluster
        workloads.water_cell(40, seed=3),  # 120 atoms of liquid water, periodic
    )):
        reqs.append(dict(rid=rid, force_field="ffield.reax.HO.2015", symbols=list(s), positions=x, cell=cell,
                         pbc=(cell is not None,) * 3, total_charge=0.0, md_steps=50, dt=0.25, temperature=300.0,
                         velocity_seed=11 + rid))
    half = []
    for r in reqs:
        h = copy.deepcopy(r)
        h["rid"] = r["rid"] + 10**6
        h["dt"] = r["dt"] / 2
        h["md_steps"] = 2 * r["md_steps"]
        half.append(h)
    res, _ = _serve(reqs + half, full_derivative=True, tol=1e-12)
    for req, h in zip(reqs, half):
        r, rh = res[req["rid"]], res[h["rid"]]
        total, total_h = r.potential + r.kinetic, rh.potential + rh.kinetic
        spread, spread_h = np.ptp(total), np.ptp(total_h)
        print(f"[serving] exact-gradient NVE rid={req['rid']} n={r.atoms} spread={spread:.3e} "
              f"(half step {spread_h:.3e}) mean KE={r.kinetic.mean():.3e}", flush=True)
        assert r.evaluations == 51 and rh.evaluations == 101
        # the lattice-built liquid cell heats from 300 K to well above it as its contacts relax
        assert spread < 1e-2 * r.kinetic.max()
        assert 0.15 < spread_h / spread < 0.32  # O(dt^2): a quarter, up to higher-order terms


def test_extrapolated_qeq_guesses_keep_results_and_cut_iterations():
    """NVE segments started from extrapolated QEq guesses give the results of the plain warm start, to
    the solve's tolerance (the server's default 1e-10 relative residual), with clearly fewer CG
    iterations (a guess closer by a fixed factor saves a fixed number of iterations, so the saving
    is larger at looser tolerances)."""
    reqs = []
    for rid, (s, x, cell) in enumerate((
        (*workloads.water_cluster_of(40, 5), None),
        workloads.water_cell(50, seed=4),
    )):
        reqs.append(dict(rid=rid, force_field="ffield.reax.HO.2015", symbols=list(s), positions=x, cell=cell,
                         pbc=(cell is not None,) * 3, total_charge=0.0, md_steps=30, dt=0.25, temperature=300.0,
                         velocity_seed=5 + rid))
    iters, out = {}, {}
    for extrapolate in (False, True):
        from xreac_amd.serving import Server

        # unrecorded steps: the count wraps evaluate_batch, which recorded replays do not call
        srv = Server(extrapolate=extrapolate, tol=1e-10, recorded_steps=False)
        count = [0]
        for eng in srv.engines.values():
            run = eng.evaluate_batch

            def counted(*a, _run=run, **k):
                res = _run(*a, **k)
                count[0] += int(np.sum(res["qeq_iterations"]))
                return res

            eng.evaluate_batch = counted
        out[extrapolate] = {r.rid: r for r in srv.run(reqs)}
        iters[extrapolate] = count[0]
    print(f"[serving] CG iterations over the segments: plain warm start {iters[False]}, "
          f"extrapolated {iters[True]}", flush=True)
    assert iters[True] < 0.85 * iters[False]
    for req in reqs:
        a, b = out[False][req["rid"]], out[True][req["rid"]]
        assert a.evaluations == b.evaluations == 31
        np.testing.assert_allclose(b.positions, a.positions, rtol=0, atol=1e-8)
        np.testing.assert_allclose(b.potential, a.potential, rtol=0, atol=1e-6)
        np.testing.assert_allclose(b.charges, a.charges, rtol=0, atol=1e-8)


def test_served_nve_with_kept_skin_lists_is_bitwise_fresh():
    """NVE segments served with kept skin lists (neighbor_skin) give bitwise the results of fresh lists,
    with far fewer list builds than evaluations."""
    reqs = []
    for rid, (s, x, cell) in enumerate((
        (*workloads.water_cluster_of(40, 5), None),
        workloads.water_cell(50, seed=4),
        (*workloads.water_cluster_of(20, 9), None),
    )):
        reqs.append(dict(rid=rid, force_field="ffield.reax.HO.2015", symbols=list(s), positions=x, cell=cell,
                         pbc=(cell is not None,) * 3, total_charge=0.0, md_steps=30, dt=0.25, temperature=300.0,
                         velocity_seed=7 + rid))
    fresh, _ = _serve(reqs)
    kept, srv = _serve(reqs, neighbor_skin=0.3)
    builds = sum(eng.neighbor_list_builds for eng in srv.engines.values())
    print(f"[serving] kept skin lists: {builds} list builds for {3 * 31} system evaluations", flush=True)
    assert builds < 31
    for req in reqs:
        a, b = fresh[req["rid"]], kept[req["rid"]]
        np.testing.assert_array_equal(a.positions, b.positions)
        np.testing.assert_array_equal(a.velocities, b.velocities)
        np.testing.assert_array_equal(a.potential, b.potential)
        np.testing.assert_array_equal(a.charges, b.charges)


def test_served_mix_with_kept_skin_lists_is_bitwise_fresh():
    """A mixed stream (single evaluations, NVE segments, relaxations) whose batches change at every step:
    kept skin lists are rebuilt whenever a group's batch changes, and every result stays bitwise that of
    fresh lists."""
    reqs = _requests(16, seed=23, relax_every=4)
    kw = dict(max_batch_atoms=500, relax_iterations=6)
    fresh, _ = _serve(reqs, **kw)
    kept, srv = _serve(reqs, neighbor_skin=0.3, **kw)
    builds = sum(eng.neighbor_list_builds for eng in srv.engines.values())
    print(f"[serving] mixed stream with kept skin lists: {srv.steps} steps, {builds} list builds", flush=True)
    assert builds > 0
    for req in reqs:
        a, b = fresh[req["rid"]], kept[req["rid"]]
        assert a.evaluations == b.evaluations
        np.testing.assert_array_equal(a.forces, b.forces)
        np.testing.assert_array_equal(a.charges, b.charges)
        np.testing.assert_array_equal(a.components, b.components)
        if req["md_steps"] or req.get("relax"):
            np.testing.assert_array_equal(a.positions, b.positions)


def test_served_relaxation_follows_oracle_fire():
    reqs = [r for r in _requests(20, seed=13, max_atoms=40) if r["md_steps"] == 0][:2]
    for r in reqs:
        r["relax"] = True
    res, _ = _serve(reqs, relax_iterations=10, tol=1e-12)
    for req in reqs:
        ff = ForceField.bundled(req["force_field"])

        def forces(x):
            return orc.evaluate(ff, req["symbols"], x, cell=req["cell"], pbc=req["pbc"],
                                total_charge=req["total_charge"])["forces"]

        x, converged, it, _ = od.fire(forces, req["positions"], 1e-4, 10)
        r = res[req["rid"]]
        print(f"[serving] FIRE rid={req['rid']} n={r.atoms} iterations={r.evaluations - 1} "
              f"converged={r.converged}", flush=True)
        assert r.converged == converged and r.evaluations == it + 1
        np.testing.assert_allclose(r.positions, x, rtol=0, atol=1e-8)


def test_recorded_steps_change_no_result():
    # NVE segments and relaxations admitted together: the group keeps its
    # composition for most steps, which then run as recorded graph replays
    reqs = [r for r in _requests(40, seed=17, md_steps=12, max_atoms=150) if r["md_steps"]][:4]
    relax = [r for r in _requests(20, seed=18, max_atoms=60) if r["md_steps"] == 0][:2]
    for r in relax:
        r["relax"] = True
    reqs = reqs + relax
    for k, r in enumerate(reqs):  # the two plans number their requests from 0 each
        r["rid"] = k
    print(f"[serving] recorded steps: {len(reqs)} requests", flush=True)
    out = {}
    for rec in (True, False):
        for precision in ("fp64", "mixed"):
            res, srv = _serve(reqs, precision=precision, relax_iterations=12, recorded_steps=rec)
            out[rec, precision] = res
            print(f"[serving] recorded={rec} {precision}: steps {srv.steps}, recorded runs {srv.recorded_runs}",
                  flush=True)
            if rec:
                assert srv.recorded_runs >= 8, srv.recorded_runs
            else:
                assert srv.recorded_runs == 0
    for precision in ("fp64", "mixed"):
        for req in reqs:
            a, b = out[True, precision][req["rid"]], out[False, precision][req["rid"]]
            assert a.error is None and b.error is None, (a.error, b.error)
            assert a.evaluations == b.evaluations
            np.testing.assert_array_equal(a.forces, b.forces)
            np.testing.assert_array_equal(a.charges, b.charges)
            np.testing.assert_array_equal(a.positions, b.positions)
            np.testing.assert_array_equal(a.potential, b.potential)
            if req["md_steps"]:
                np.testing.assert_array_equal(a.velocities, b.velocities)
                np.testing.assert_array_equal(a.kinetic, b.kinetic)
