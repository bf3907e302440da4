"""Batched multi-scene engine against the plain per-scene path.

Each scene advanced in a batch must give what the WCSPH solver gives for
that scene alone on the same grid (fields within 1e-6 relative; the engine
is built to give the same bits), and no schedule, autotuner choice or graph
replay may change any result.
"""
import numpy as np
import pytest
import torch

import pysph_amd.base.api as base
import pysph_amd.solver.api as solver
from pysph_amd.batch.engine import BatchedWCSPH, scene_from_arrays
from pysph_amd.solver import generators as gen

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "u", "v", "rho", "p")


def _scenes(n, seed, n_min=600, n_max=5000):
    return gen.batch_scenes(n, seed=seed, n_min=n_min, n_max=n_max)


def _plain(f, b, info, nsteps, cells_per_radius=2):
    arrays = [f] + ([b] if b is not None else [])
    particles = base.Particles(arrays, update_particles=False)
    cm = particles.cell_manager
    lo, hi = info["domain"]
    cm.set_fixed_domain(lo, hi)
    cm.check_jumps = False
    s = solver.wcsph_solver(2, integrator="rk2", co=info["co"], ro=info["rho0"], gamma=7.0,
                            alpha=info["alpha"], beta=info["beta"], eta=info["eta"],
                            gravity=info["g"], xsph_eps=info["eps"] or None,
                            with_solid=b is not None)
    s.setup_integrator(particles)
    plan = s.integrator.plan
    assert plan is not None and plan.use_packed
    plan.set_cells_per_radius(cells_per_radius)
    for _ in range(nsteps):
        s.integrator.integrate(info["dt"])
    return {c: f.get(c) for c in FIELDS}, (None if b is None else b.get("rho"))


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30))) if len(a) else 0.0


def test_batch_matches_plain_per_scene():
    nsteps = 10
    sc = _scenes(6, seed=3)
    batch = BatchedWCSPH([scene_from_arrays(f, b, i) for f, b, i in sc])
    batch.run(nsteps)
    torch.cuda.synchronize()
    bitwise = True
    for k, (f, b, info) in enumerate(_scenes(6, seed=3)):
        ref, rho_b = _plain(f, b, info, nsteps)
        got = batch.scene_state(k, FIELDS)
        for c in FIELDS:
            scale = np.max(np.abs(ref[c])) if c in ("x", "y", "u", "v") else None
            if scale:
                err = float(np.max(np.abs(got[c] - ref[c]))) / scale
            else:
                err = _rel(got[c], ref[c])
            assert err <= 1e-6, (k, info["kind"], c, err)
            bitwise &= bool(np.array_equal(got[c], ref[c]))
        if rho_b is not None:
            assert _rel(batch.boundary_state(k, ("rho",))["rho"], rho_b) <= 1e-6
    print("batch == plain bitwise:", bitwise)
    assert bitwise


def _final(batch, nsteps):
    batch.run(nsteps)
    torch.cuda.synchronize()
    return [batch.scene_state(k, FIELDS) for k in range(len(batch.scenes))]


@pytest.mark.parametrize("sched", [
    {"mode": 1, "blocks_per_cu": 2, "grab": 1, "order": 0},
    {"mode": 1, "blocks_per_cu": 8, "grab": 1, "order": 1},
    {"mode": 2, "blocks_per_cu": 4, "grab": 1, "order": 0},
    {"mode": 2, "blocks_per_cu": 4, "grab": 4, "order": 1}])
def test_schedules_give_identical_bits(sched):
    sc = _scenes(12, seed=5)
    ref = _final(BatchedWCSPH([scene_from_arrays(f, b, i) for f, b, i in sc]), 3)
    eng = BatchedWCSPH([scene_from_arrays(f, b, i) for f, b, i in sc])
    eng.sched = {"density": dict(sched), "stage": dict(sched)}
    got = _final(eng, 3)
    for r, g in zip(ref, got):
        for c in FIELDS:
            assert np.array_equal(r[c], g[c]), (sched, c)


def test_autotune_then_graph_replays_match_eager():
    sc = _scenes(10, seed=7)
    eager = BatchedWCSPH([scene_from_arrays(f, b, i) for f, b, i in sc])
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        eng = BatchedWCSPH([scene_from_arrays(f, b, i) for f, b, i in sc])
        res = eng.autotune(reps=2)
        assert set(res) == {"density", "stage"}
        eng.capture()
        for rep in range(4):
            eng.step()
            torch.cuda.synchronize()
            eager.step()
            torch.cuda.synchronize()
            # eager host work between replays
            _ = torch.ones(1000, device="cuda").sum().item()
            for k in range(len(sc)):
                a, b_ = eager.scene_state(k, FIELDS), eng.scene_state(k, FIELDS)
                for c in FIELDS:
                    assert np.array_equal(a[c], b_[c]), (rep, k, c)
        eng.close()


def _oracle_scene(f, b, info, nsteps):
    """The FP64 oracle's RK2 steps (integrator.py:581-645 per scene) from the
    scene's stored FP32 bits."""
    from oracle import oracle as O
    from tests.util import to_oracle
    arrs = [to_oracle(f, 0)] + ([to_oracle(b, 1)] if b is not None else [])
    sch = O.scheme(2, integrator="rk2", dt=info["dt"], co=info["co"], ro=info["rho0"],
                   gamma=7.0, alpha=info["alpha"], beta=info["beta"], eta=info["eta"],
                   gravity=info["g"], xsph_eps=info["eps"] or None, cell_size=2.0 * info["h"])
    O.scheme_step(sch, arrs, nsteps)
    return arrs


def _check_scene_vs_oracle(batch, k, f, b, info, nsteps):
    from tests.util import max_rel, rel_norm
    ref = _oracle_scene(f, b, info, nsteps)
    got = batch.scene_state(k, FIELDS)
    scale = max(1.0, float(np.max(np.abs(ref[0]["x"]))), float(np.max(np.abs(ref[0]["y"]))))
    for c in ("x", "y"):
        # FP32 storage of the coordinates: a few units in the last place
        assert np.max(np.abs(got[c] - ref[0][c])) < 5e-7 * scale, (k, info["kind"], c)
    ev = rel_norm(np.stack([got["u"], got["v"]]), np.stack([ref[0]["u"], ref[0]["v"]]))
    er = max_rel(got["rho"], ref[0]["rho"])
    assert ev < 1e-5 and er < 1e-5, (k, info["kind"], ev, er)
    if b is not None:
        assert max_rel(batch.boundary_state(k, ("rho",))["rho"], ref[1]["rho"]) < 1e-5
    return ev, er


def test_batch_scenes_vs_oracle():
    """Scenes of one batch (elliptical drops and dam breaks of 1k-16k
    particles) against the FP64 oracle, 3 RK2 steps from the same stored
    bits: density max relative and velocity normwise within 1e-5."""
    nsteps = 3
    sc = _scenes(8, seed=11, n_min=1000, n_max=16000)
    assert {i["kind"] for _, _, i in sc} == {"drop", "dam"}
    batch = BatchedWCSPH([scene_from_arrays(f, b, i) for f, b, i in sc])
    batch.run(nsteps)
    torch.cuda.synchronize()
    for k, (f, b, info) in enumerate(sc):
        ev, er = _check_scene_vs_oracle(batch, k, f, b, info, nsteps)
        print("scene %d (%s, %d fluid): velocity %.2e, rho %.2e"
              % (k, info["kind"], f.get_number_of_particles(), ev, er))


def test_full_size_batch_vs_oracle():
    """The benchmarked batch itself (512 scenes, 1k-64k fluid particles
    each, one HIP graph per step, autotuned schedules), 3 RK2 steps; the
    largest and the smallest scene of each kind against the FP64 oracle."""
    from oracle import oracle as O
    nsteps = 3
    sc = gen.batch_scenes(512, seed=0)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        eng = BatchedWCSPH([scene_from_arrays(f, b, i) for f, b, i in sc])
        eng.autotune(reps=1)
        eng.capture()
        for _ in range(nsteps):
            eng.step()
        torch.cuda.synchronize()
    picks = []
    for kind in ("drop", "dam"):
        ks = [k for k, (_, _, i) in enumerate(sc) if i["kind"] == kind]
        ks.sort(key=lambda k: sc[k][0].get_number_of_particles())
        picks += [ks[0], ks[-1]]
    O.set_threads(0)
    try:
        for k in picks:
            f, b, info = sc[k]
            ev, er = _check_scene_vs_oracle(eng, k, f, b, info, nsteps)
            print("scene %d (%s, %d fluid): velocity %.2e, rho %.2e"
                  % (k, info["kind"], f.get_number_of_particles(), ev, er))
    finally:
        O.set_threads(1)
        eng.close()
