"""The five integrators on the reference's own circle test
(solver/tests/test_integrator.py:1-385): four particles on a circle of
radius 2/pi moved by MoveCircleX / MoveCircleY (unit speed) for one second
end a quarter turn later, to the reference's number of decimal places."""
import numpy as np
import pytest

import pysph_amd.base.api as base
import pysph_amd.solver.api as solver
import pysph_amd.sph.api as sph

pytestmark = pytest.mark.gpu

Fluids = base.ParticleType.Fluid


def _setup(cls):
    r = 2.0 / np.pi
    x = np.array([1.0, 0.0, -1.0, 0.0]) * r
    y = np.array([0.0, 1.0, 0.0, -1.0]) * r
    pa = base.get_particle_array(x=x, y=y, p=np.zeros(4), e=np.zeros(4), h=np.ones(4),
                                 name="tmp")
    particles = base.Particles(arrays=[pa])
    kernel = base.CubicSplineKernel(dim=2)
    circlex = solver.SPHIntegration(sph.MoveCircleX, from_types=[Fluids], on_types=[Fluids],
                                    updates=["x", "y"], id="circlex")
    circley = solver.SPHIntegration(sph.MoveCircleY, from_types=[Fluids], on_types=[Fluids],
                                    updates=["x", "y"], id="circley")
    calcs = circlex.get_calcs(particles, kernel) + circley.get_calcs(particles, kernel)
    integrator = cls(particles=particles, calcs=calcs)
    return integrator, particles, pa, r


@pytest.mark.parametrize("cls,nsteps,places", [
    (solver.EulerIntegrator, 1, 2),
    (solver.RK2Integrator, 2, 3),
    (solver.RK4Integrator, 4, 12),
    (solver.PredictorCorrectorIntegrator, 2, 6),
    (solver.LeapFrogIntegrator, 2, 2)])
def test_circle_motion(cls, nsteps, places):
    integrator, particles, pa, r = _setup(cls)
    assert integrator.nsteps == nsteps
    integrator.setup_integrator()
    assert integrator.plan is None
    t, tf, dt = 0.0, 1.0, 1e-3
    while t <= tf:
        t += dt
        particles.update()
        integrator.integrate(dt)
    exact = np.array([(0.0, r), (-r, 0.0), (0.0, -r), (r, 0.0)])
    got = np.stack([pa.get("x"), pa.get("y")], 1)
    # assertAlmostEqual(a, b, places): round(a - b, places) == 0
    assert np.all(np.round(got - exact, places) == 0), (got, exact)


@pytest.mark.parametrize("cls", [solver.RK4Integrator, solver.PredictorCorrectorIntegrator,
                                 solver.LeapFrogIntegrator])
def test_wcsph_dam_break_runs_with_each_integrator(cls):
    """The WCSPH op set through the generic calc path with the other
    integrators: five steps stay finite and conserve mass."""
    from pysph_amd.solver import generators as gen
    f, b, info = gen.dam_break_2d(width=0.2, height=0.3, dx=0.01, jitter=0.05, seed=2,
                                  precision="double")
    particles = base.Particles([f, b])
    s = solver.wcsph_solver(2, co=info["co"], ro=info["rho0"], alpha=0.1, gravity=info["g"],
                            xsph_eps=0.5)
    s.integrator_type = cls
    s.setup_integrator(particles)
    assert isinstance(s.integrator, cls) and s.integrator.plan is None
    m0 = f.get("m").sum()
    y0 = f.get("y").mean()
    for _ in range(5):
        s.integrator.integrate(info["dt"])
    for c in ("x", "y", "u", "v", "rho"):
        assert np.all(np.isfinite(f.get(c)))
    assert f.get("m").sum() == m0
    assert f.get("y").mean() < y0  # the column falls
