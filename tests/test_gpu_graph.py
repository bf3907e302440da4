"""HIP graph replay of the fused WCSPH RK2 step (WCSPHPlan.step): replayed
steps give the same bits as the eagerly launched steps on a fixed domain,
stay within FP32 summation-order differences of them when the grid follows
the particle bounds (and has to change), and a default drop-in solver
replays graphs.  Each replay is followed by the program's own eager work and
a synchronisation (host reads of every field)."""
import numpy as np
import pytest
import torch

import pysph_amd.base.api as base
import pysph_amd.solver.api as solver
from pysph_amd.solver import generators as gen

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "u", "v", "w", "rho", "p", "cs")


def _dam(use_graph, fixed, check_jumps="deferred"):
    f, b, info = gen.dam_break_3d(dx=1.0 / 40, column=(10, 10, 20), tank=(24, 10, 20), seed=5)
    particles = base.Particles([f, b], update_particles=False)
    cm = particles.cell_manager
    if fixed:
        cm.set_fixed_domain(*info["domain"])
    cm.check_jumps = check_jumps
    s = solver.wcsph_solver(3, integrator="rk2", co=info["co"], ro=info["rho0"], gamma=7.0,
                            alpha=info["alpha"], beta=info["beta"], eta=info["eta"],
                            gravity=info["g"], xsph_eps=info["eps"])
    s.setup_integrator(particles)
    s.set_time_step(info["dt"])
    plan = s.integrator.plan
    assert plan is not None and plan.graph_applies()
    plan.use_graph = use_graph
    return s, f, b, info


def _state(*arrays):
    return {(pa.name, k): pa.get(k) for pa in arrays for k in FIELDS}


def test_graph_replays_give_the_eager_bits_on_a_fixed_domain():
    se, fe, be, info = _dam(False, True)
    sg, fg, bg, _ = _dam(True, True)
    for k in range(14):
        se.integrator.integrate(info["dt"])
        sg.integrator.integrate(info["dt"])
        torch.cuda.synchronize()
        a, b = _state(fe, be), _state(fg, bg)
        for key in a:
            assert np.array_equal(a[key], b[key]), (k, key)
    plan = sg.integrator.plan
    assert plan.graph_error is None, plan.graph_error
    # steps 1-3 eager (no grid before the first, then two warm-up steps),
    # one capture, replays from step 4 on
    assert plan.graph_captures == 1 and plan.graph_replays == 11
    assert se.integrator.plan.graph_replays == 0
    se.particles.cell_manager.verify_jumps()
    sg.particles.cell_manager.verify_jumps()


def test_graph_replays_follow_a_grid_change():
    """Without a fixed domain the grid follows bounds read back without a
    wait.  Halfway the fluid is lifted above the grid: the replays go on
    (border cells hold it exactly) until the bounds reach the host, then
    the plan steps eagerly on a new grid and captures again."""
    se, fe, be, info = _dam(False, False, check_jumps=False)
    sg, fg, bg, _ = _dam(True, False, check_jumps=False)
    for k in range(24):
        if k == 10:
            for f in (fe, fg):
                f.set(z=f.get("z") + 0.4)
        se.integrator.integrate(info["dt"])
        sg.integrator.integrate(info["dt"])
        torch.cuda.synchronize()
        a, b = _state(fe, be), _state(fg, bg)
        for key in a:
            ref = np.abs(a[key]).max() + 1e-30
            assert np.abs(a[key] - b[key]).max() <= 1e-5 * ref, (k, key)
    plan = sg.integrator.plan
    assert plan.graph_error is None, plan.graph_error
    assert plan.graph_captures >= 2 and plan.graph_replays >= 12
    assert sg.particles.cell_manager.grid is not None
    g = sg.particles.cell_manager.grid
    z = fg.get("z")
    assert g.lo[2] * g.cell_size <= z.min() and z.max() <= (g.lo[2] + g.dims[2]) * g.cell_size


def test_default_solver_replays_graphs_and_checks_jumps():
    s, f, b, info = _dam(True, False, check_jumps=True)
    s.set_final_time(9.5 * info["dt"])
    s.set_print_freq(0)
    s.solve()
    plan = s.integrator.plan
    assert s.count == 10
    assert plan.graph_error is None and plan.graph_replays >= 7
    assert np.all(np.isfinite(f.get("x")))


def test_jump_check_raises_under_graph_replay():
    """The jump check of a replayed step counts on the device inside the
    graph; the count is read back between replays and raises the reference's
    JumpToleranceError (cell.pyx:1422-1437) at a following step or at the
    end of solve()."""
    from pysph_amd._abi import JumpToleranceError
    s, f, b, info = _dam(True, True, check_jumps=True)
    for _ in range(6):
        s.integrator.integrate(info["dt"])
    plan = s.integrator.plan
    assert plan.graph_replays >= 2 and plan.graph_error is None
    z = f.get("z")
    z[0] = z[0] + 0.3  # 0.3 m: many cells of 2h = 0.06 m
    f.set(z=z)
    with pytest.raises(JumpToleranceError):
        for _ in range(4):
            s.integrator.integrate(info["dt"])
        torch.cuda.synchronize()
        s.particles.cell_manager.verify_jumps()


@pytest.mark.parametrize("integrator", ["rk2", "euler"])
def test_graph_replays_of_the_fp64_soa_step(integrator):
    """FP64 storage (the reference's default precision) runs the SoA
    kernels, re-binning by the predicted positions in RK2 stage 2 (X1 and
    k1 alternate between two buffer sets) and swapping the state with the
    stage output in Euler: the buffers come back after two (RK2) or six
    (Euler: the swapped fields cycle through three buffers, the others
    through two) steps, and every replayed step gives the eager bits."""
    runs = []
    for use_graph in (False, True):
        pa = gen.elliptical_drop(dx=0.05, precision="double")
        particles = base.Particles([pa])
        s = solver.wcsph_solver(2, integrator=integrator, co=1400.0, ro=1.0, gamma=7.0,
                                alpha=0.01, beta=0.0, eta=0.1, with_solid=False)
        s.setup_integrator(particles)
        plan = s.integrator.plan
        assert plan is not None and not plan.use_packed
        plan.use_graph = use_graph
        runs.append((s, pa, plan))
    for k in range(12):
        for s, pa, plan in runs:
            s.integrator.integrate(1e-5)
        torch.cuda.synchronize()
        a, b = _state(runs[0][1]), _state(runs[1][1])
        for key in a:
            assert np.array_equal(a[key], b[key]), (k, key)
    plan = runs[1][2]
    assert plan.graph_error is None, plan.graph_error
    assert plan.graph_captures == 1 and plan.graph_replays >= 8
    assert len(plan._sg._g["exec"]) == (2 if integrator == "rk2" else 6)


def test_graph_replays_after_particles_are_added():
    """Particles added between steps give the arrays new buffers: the plan
    steps eagerly, captures again, and every step keeps the eager bits."""
    runs = [_dam(False, True), _dam(True, True)]
    for k in range(16):
        if k == 7:
            for s, f, b, info in runs:
                x = f.get("x")
                sel = slice(0, 40)
                # a layer dropped above the column (and above the fixed
                # domain: clamped into its border bins)
                f.add_particles(x=x[sel], y=f.get("y")[sel], z=f.get("z")[sel] + 0.6,
                                h=f.get("h")[sel], m=f.get("m")[sel], rho=f.get("rho")[sel])
        for s, f, b, info in runs:
            s.integrator.integrate(info["dt"])
        torch.cuda.synchronize()
        a, b2 = _state(runs[0][1], runs[0][2]), _state(runs[1][1], runs[1][2])
        for key in a:
            assert np.array_equal(a[key], b2[key]), (k, key)
    plan = runs[1][0].integrator.plan
    assert plan.graph_error is None, plan.graph_error
    assert plan.graph_captures == 2
    assert runs[1][1].get_number_of_particles() == runs[0][1].get_number_of_particles()
