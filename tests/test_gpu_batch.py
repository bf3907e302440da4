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


def _orc_scheme(info):
    from oracle import oracle as O
    return O.scheme(2, integrator="rk2", dt=info["dt"], co=info["co"], ro=info["rho0"],
                    gamma=7.0, alpha=info["alpha"], beta=info["beta"], eta=info["eta"],
                    gravity=info["g"], xsph_eps=info["eps"] or None, cell_size=2.0 * info["h"])


def _stage_state(eng, k):
    """X1 and k1 of scene k after a step (the engine's stage buffers, in the
    step's binning order fb), in the scene's user order."""
    a, b = int(eng.f_off[k]), int(eng.f_off[k + 1])
    uid = eng.fb["uid"][a:b].cpu().numpy()

    def user(t):
        v = t[a:b].cpu().numpy().astype(np.float64)
        r = np.empty_like(v)
        r[uid] = v
        return r
    return [user(t) for t in eng.x1a], [user(t) for t in eng.k1a]


def _step_vs_oracle(eng, k, f, b, info):
    """One RK2 step of scene k (the engine has just taken it) against the
    FP64 oracle on the same stored inputs, stage by stage: k1 = (a, dv) at
    the step's start state X, the density at the engine's predictor X1, and
    the final velocity u0 + dt/2 (k1 + k2) + dt g with k2 the oracle's at X1
    (integrator.py:581-645, pressure_funcs.pyx:154-245, xsph_funcs.pyx:31-88,
    density_funcs.pyx:38-84).  Returns the errors."""
    from oracle import oracle as O
    from tests.util import max_rel, rel_norm
    x0 = {c: f[c] for c in ("x", "y", "u", "v")}
    sch = _orc_scheme(info)
    rf = O.make_array(ptype=0, **{c: f[c] for c in f})
    arrs = [rf] + ([O.make_array(ptype=1, **{c: b[c] for c in b})] if b is not None else [])
    a_o, d_o = O.scheme_rhs(sch, arrs)
    X1, K1 = _stage_state(eng, k)
    ek1 = rel_norm(np.stack([K1[0], K1[1]]), a_o[:2])
    ed1 = rel_norm(np.stack([K1[3], K1[4]]), d_o[:2]) if info["eps"] else 0.0
    r1 = dict(rf)
    for c, nm in ((0, "x"), (1, "y"), (3, "u"), (4, "v")):
        r1[nm] = X1[c]
    arrs1 = [O.make_array(ptype=0, **{c: r1[c] for c in f})] + arrs[1:]
    a_o2, d_o2 = O.scheme_rhs(sch, arrs1)
    got = eng.scene_state(k, ("x", "y", "u", "v", "rho"))
    erho = max_rel(got["rho"], arrs1[0]["rho"])
    hdt = 0.5 * info["dt"]
    g = info["g"] or (0.0, 0.0, 0.0)
    u_o = [x0[c] + hdt * (K1[i] + a_o2[i]) + hdt * (g[i] + g[i])
           for i, c in ((0, "u"), (1, "v"))]
    eu = rel_norm(np.stack([got["u"], got["v"]]), np.stack(u_o))
    return ek1, ed1, erho, eu


def _snapshot(eng, k, with_boundary):
    st = eng.scene_state(k, ("x", "y", "z", "u", "v", "w", "h", "m", "rho", "p", "cs"))
    f = {c: v.astype(np.float64) for c, v in st.items()}
    b = None
    if with_boundary:
        bs = eng.boundary_state(k, ("x", "y", "z", "u", "v", "w", "h", "m", "rho", "p", "cs"))
        b = {c: v.astype(np.float64) for c, v in bs.items()}
    return f, b


def _check_steps(eng, ks, sc, nsteps, label):
    worst = 0.0
    for step in range(nsteps):
        snaps = {k: _snapshot(eng, k, sc[k][1] is not None) for k in ks}
        eng.step()
        torch.cuda.synchronize()
        for k in ks:
            f, b = snaps[k]
            info = sc[k][2]
            errs = _step_vs_oracle(eng, k, f, b, info)
            print("%s step %d scene %d (%s, %d fluid): k1 %.2e, xsph %.2e, rho(X1) %.2e, "
                  "u %.2e" % ((label, step, k, info["kind"], len(f["x"])) + errs))
            assert max(errs) < 1e-5, (step, k, info["kind"], errs)
            worst = max(worst, max(errs))
    return worst


def test_batch_scenes_vs_oracle():
    """Scenes of one batch (elliptical drops and dam breaks of 1k-16k
    particles), 3 RK2 steps; every step checked against the FP64 oracle on
    the engine's own stored inputs: k1, the density at X1 and the final
    velocity within 1e-5 (density max relative, the rest normwise)."""
    sc = _scenes(8, seed=11, n_min=1000, n_max=16000)
    assert {i["kind"] for _, _, i in sc} == {"drop", "dam"}
    eng = BatchedWCSPH([scene_from_arrays(f, b, i) for f, b, i in sc])
    _check_steps(eng, list(range(len(sc))), sc, 3, "batch")


def test_full_size_batch_vs_oracle():
    """The benchmarked batch itself (512 scenes, 1k-64k fluid particles
    each, one HIP graph per step, autotuned schedules), 3 RK2 steps; the
    largest and the smallest scene of each kind checked every step against
    the FP64 oracle on the engine's stored inputs, as above."""
    from oracle import oracle as O
    sc = gen.batch_scenes(512, seed=0)
    picks = []
    for kind in ("drop", "dam"):
        ks = [k for k, (_, _, i) in enumerate(sc) if i["kind"] == kind]
        ks.sort(key=lambda k: sc[k][0].get_number_of_particles())
        picks += [ks[0], ks[-1]]
    stream = torch.cuda.Stream()
    O.set_threads(0)
    try:
        with torch.cuda.stream(stream):
            eng = BatchedWCSPH([scene_from_arrays(f, b, i) for f, b, i in sc])
            eng.autotune(reps=1)
            eng.capture()
            _check_steps(eng, picks, sc, 3, "full batch")
            eng.close()
    finally:
        O.set_threads(1)
