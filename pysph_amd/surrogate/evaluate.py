"""Surrogate against the plain solver on one 2D dam break: the reference is
the plain step at dt, the method without the network is the plain step at
dt_factor x dt."""
import math
import time

import numpy as np
import torch

from .. import _abi
from .graph import ParticleGraph, combined_array
from .solver import STATE, SurrogateWCSPH, get_user, plain_solver, set_user
from .train import dam_scene


def _density(fluid, boundary, info):
    g = ParticleGraph(combined_array(fluid, boundary), info, domain=info["domain"])
    g.build()
    return get_user(g.pa, "rho")[:fluid.get_number_of_particles()]


def compare(model, seed, nsteps=100, n=30000, dt_factor=4, check_every=10, threshold=0.01,
            warm=0, log=None, baseline=None):
    """Run the three methods from the same initial state for
    nsteps * dt_factor plain steps (after `warm` plain steps of set-up) and
    return errors and timings."""
    def scene():
        rng = np.random.Generator(np.random.PCG64(seed))
        f, b, info = dam_scene(rng, (n, n))
        if warm:
            s = plain_solver(f, b, info)
            for _ in range(warm):
                s.integrator.integrate(info["dt"])
        return f, b, info

    out = {"seed": seed, "steps": nsteps, "dt_factor": dt_factor}
    # reference; both timed runs start warm: a few untimed steps (first-call
    # allocations, kernel loading) and a restart from the saved state
    f, b, info = scene()
    s = plain_solver(f, b, info)
    snap = {k: get_user(f, k) for k in STATE}
    for _ in range(2 * dt_factor):
        s.integrator.integrate(info["dt"])
    for k in STATE:
        set_user(f, k, snap[k])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsteps * dt_factor):
        s.integrator.integrate(info["dt"])
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t0
    rho_ref = _density(f, b, info)
    x_ref = torch.stack([get_user(f, "x"), get_user(f, "y")])
    span = nsteps * dt_factor * info["dt"]
    out.update(n_fluid=f.get_number_of_particles(), n_boundary=b.get_number_of_particles(),
               dt=info["dt"], simulated_time=span,
               plain_sim_time_per_s=span / t_ref, plain_seconds=t_ref)
    # surrogate
    f2, b2, info2 = scene()
    sur = SurrogateWCSPH(f2, b2, info2, model, dt_factor=dt_factor, check_every=check_every,
                         threshold=threshold)
    st0 = sur.state()
    sur.run(check_every)
    sur.reset(st0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sur.run(nsteps)
    torch.cuda.synchronize()
    t_sur = time.perf_counter() - t0
    st = sur.state()
    nf = f2.get_number_of_particles()
    rho_sur = sur.density()[:nf]
    x_sur = torch.stack([st["x"][:nf], st["y"][:nf]])
    out.update(surrogate_sim_time_per_s=span / t_sur, surrogate_seconds=t_sur,
               surrogate_density_error=float(((rho_sur - rho_ref).abs().mean() /
                                              info["rho0"]).item()),
               surrogate_position_error=float(((x_sur - x_ref).norm(dim=0).mean() /
                                               info["dx"]).item()),
               checks=sur.stats["checks"], fallbacks=sur.stats["fallbacks"],
               check_errors=[round(e, 6) for e in sur.stats["errors"]])
    # the same surrogate stepping without the learned correction
    if baseline is not None:
        f4, b4, info4 = scene()
        base = SurrogateWCSPH(f4, b4, info4, baseline, dt_factor=dt_factor,
                              check_every=check_every, threshold=threshold)
        base.run(nsteps)
        stb = base.state()
        rho_b = base.density()[:nf]
        x_b = torch.stack([stb["x"][:nf], stb["y"][:nf]])
        out.update(untrained_density_error=float(((rho_b - rho_ref).abs().mean() /
                                                  info["rho0"]).item()),
                   untrained_position_error=float(((x_b - x_ref).norm(dim=0).mean() /
                                                   info["dx"]).item()),
                   untrained_fallbacks=base.stats["fallbacks"],
                   untrained_check_errors=[round(e, 6) for e in base.stats["errors"]])
    # plain at dt_factor x dt
    f3, b3, info3 = scene()
    s3 = plain_solver(f3, b3, info3)
    ok = True
    try:
        for _ in range(nsteps):
            s3.integrator.integrate(info3["dt"] * dt_factor)
        torch.cuda.synchronize()
        rho_big = _density(f3, b3, info3)
        x_big = torch.stack([get_user(f3, "x"), get_user(f3, "y")])
        e_rho = float(((rho_big - rho_ref).abs().mean() / info["rho0"]).item())
        e_x = float(((x_big - x_ref).norm(dim=0).mean() / info["dx"]).item())
        if not (math.isfinite(e_rho) and math.isfinite(e_x)):
            ok = False
    except _abi.SphError as e:
        ok, e_rho, e_x = False, float("nan"), float("nan")
        out["big_dt_failure"] = str(e)[:200]
    out.update(big_dt_finite=ok, big_dt_density_error=e_rho, big_dt_position_error=e_x)
    if log:
        log(str(out))
    return out
