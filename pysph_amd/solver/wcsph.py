"""The standard weakly-compressible SPH set-up of the reference's
examples, assembled from the reference's own operation classes
(examples/dam-break/dam_break.py:270-326, examples/elliptical_drop.py,
solver/fluid_solver.py:53-83), with summation density."""
from ..base.particle_types import ParticleType
from ..sph.funcs import GravityForce, MomentumEquation, SPHRho, TaitEquation
from .integrator import EulerIntegrator, RK2Integrator
from .solver import Solver
from .sph_equation import SPHIntegration, SPHOperation

Fluid = ParticleType.Fluid
Solid = ParticleType.Solid


def wcsph_solver(dim, integrator="rk2", co=1.0, ro=1000.0, gamma=7.0, alpha=0.1, beta=0.0,
                 eta=0.1, gravity=None, xsph_eps=None, with_solid=True):
    """Solver with: density (SPHRho) and eos (TaitEquation) on fluid and
    solid, momentum (MomentumEquation) on fluid from fluid and solid,
    optional gravity, position stepping and optional XSPH on fluid."""
    itype = RK2Integrator if integrator == "rk2" else EulerIntegrator
    s = Solver(dim=dim, integrator_type=itype)
    types = [Fluid, Solid] if with_solid else [Fluid]
    vel = ["u", "v", "w"][:dim]
    s.add_operation(SPHOperation(SPHRho.withargs(), on_types=types, from_types=types,
                                 updates=["rho"], id="density"))
    s.add_operation(SPHOperation(TaitEquation.withargs(co=co, ro=ro, gamma=gamma),
                                 on_types=types, updates=["p", "cs"], id="eos"))
    s.add_operation(SPHIntegration(
        MomentumEquation.withargs(alpha=alpha, beta=beta, eta=eta, dim=dim),
        on_types=[Fluid], from_types=types, updates=vel, id="mom"))
    if gravity is not None:
        g = tuple(gravity) + (0.0,) * (3 - len(gravity))
        s.add_operation(SPHIntegration(GravityForce.withargs(gx=g[0], gy=g[1], gz=g[2]),
                                       on_types=[Fluid], updates=vel, id="gravity"))
    s.add_operation_step([Fluid])
    if xsph_eps:
        s.add_operation_xsph(eps=xsph_eps)
    return s
