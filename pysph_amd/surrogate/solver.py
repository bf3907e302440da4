"""Time stepping with the surrogate: steps of dt_factor x the plain dt,
checked against the plain WCSPH step every `check_every` steps, with the
plain step as the fallback whenever the density error exceeds `threshold`.

The surrogate steps by leapfrog (kick, then drift) with DT = dt_factor x dt:
between its steps the velocity is the mean velocity of the last step,
u_{n-1/2}.  A state that comes from the plain solver carries the
instantaneous velocity u_n instead, so the first surrogate step after it
kicks by half a step (u_{n+1/2} = u_n + DT/2 a_n), and a surrogate state
handed to the plain solver gets u_n = (u_{n-1/2} + u_{n+1/2}) / 2."""
import ctypes

import numpy as np
import torch

from .. import _abi
from .._device import require_gpu
from ..base.particles import Particles
from ..base.step_graph import StepGraphs
from ..solver.wcsph import wcsph_solver
from .graph import ParticleGraph, combined_array

STATE = ("x", "y", "u", "v")


def get_user(pa, name):
    t = pa.get_carray(name)
    o = pa._order
    if o is None:
        return t.clone()
    out = torch.empty_like(t)
    out[o] = t
    return out


def set_user(pa, name, vals):
    t = pa.get_carray(name)
    o = pa._order
    t.copy_(vals if o is None else vals[o])
    if name in ("x", "y", "z"):
        pa.set_dirty(True)


def plain_solver(fluid, boundary, info, cells_per_radius=2):
    """The plain fused WCSPH solver of a 2D dam break on a fixed domain."""
    arrays = [fluid] + ([boundary] if boundary is not None else [])
    particles = Particles(arrays, update_particles=False)
    cm = particles.cell_manager
    cm.set_fixed_domain(*info["domain"])
    cm.check_jumps = False
    s = wcsph_solver(2, integrator="rk2", co=info["co"], ro=info["rho0"], gamma=7.0,
                     alpha=info["alpha"], beta=info["beta"], eta=info["eta"], gravity=info["g"],
                     xsph_eps=info["eps"] or None, with_solid=boundary is not None)
    s.setup_integrator(particles)
    if s.integrator.plan is not None:
        s.integrator.plan.set_cells_per_radius(cells_per_radius)
    s.set_time_step(info["dt"])
    return s


class SurrogateWCSPH:
    """Surrogate stepping of one 2D dam break (fluid + fixed boundary)."""

    def __init__(self, fluid, boundary, info, model, dt_factor=4, check_every=10,
                 threshold=0.01, cells_per_radius=2):
        self.info = dict(info)
        self.model = model
        self.dt = float(info["dt"])
        self.dt_factor = int(dt_factor)
        self.check_every = int(check_every)
        self.threshold = float(threshold)
        self.nf = fluid.get_number_of_particles()
        self.graph = ParticleGraph(combined_array(fluid, boundary), info, domain=info["domain"],
                                   cells_per_radius=cells_per_radius)
        self.plain = plain_solver(fluid, boundary, info, cells_per_radius)
        self.fluid, self.boundary = fluid, boundary
        self._out = None
        self.t = 0.0
        self.steps = 0
        self.plain_left = 0
        # the state's velocity is instantaneous (from the plain solver)
        self.inst = True
        self.stats = {"surrogate": 0, "plain": 0, "checks": 0, "fallbacks": 0, "errors": []}
        # full-kick steps replay HIP graphs of graph build + network +
        # update (the build bins the graph's array, fixed domain)
        self.use_graph = True
        self.graphs = StepGraphs([self.graph.pa], self.graph.cm)

    # ---- state transfer (user order: fluid then boundary)
    def state(self):
        pa = self.graph.pa
        return {f: get_user(pa, f) for f in STATE}

    def set_state(self, st):
        for f in STATE:
            set_user(self.graph.pa, f, st[f])

    def _plain_interval(self, st):
        """dt_factor plain steps from state st; returns the new state."""
        f = self.fluid
        for k in STATE:
            set_user(f, k, st[k][:self.nf])
        for _ in range(self.dt_factor):
            self.plain.integrator.integrate(self.dt)
        out = {k: st[k].clone() for k in STATE}
        for k in STATE:
            out[k][:self.nf] = get_user(f, k)
        return out

    def _surrogate_interval(self):
        """One leapfrog step: kick by DT (half of it from an instantaneous
        velocity), then drift by DT."""
        if self.use_graph and not self.inst:
            self.graphs.step(self._surrogate_body, tag=self.dt * self.dt_factor)
        else:
            self._surrogate_body()
        self.inst = False

    def _surrogate_body(self):
        g = self.graph.build()
        n = g.n
        if self._out is None or self._out.shape[0] < n:
            self._out = torch.empty((max(n, 1), 2), dtype=torch.float32, device=g.device)
        out = self.model.forward(g, self._out, infer=True)
        pa = g.pa
        # the kernel kicks by dt * a_scale * a and drifts by dt
        a_scale = 9.81 * (0.5 if self.inst else 1.0)
        _abi.call("sph_mlp_integrate", require_gpu(out), n, pa.get_carray("btype").data_ptr(),
                  out.data_ptr(), a_scale, self.dt * self.dt_factor,
                  pa.get_carray("x").data_ptr(), pa.get_carray("y").data_ptr(),
                  pa.get_carray("u").data_ptr(), pa.get_carray("v").data_ptr())
        pa.set_dirty(True)

    def density(self, st=None):
        """Summation density (user order) of a state (default: current)."""
        if st is not None:
            self.set_state(st)
        self.graph.build(edges=False)
        rho = get_user(self.graph.pa, "rho")
        self.graph.check_overflow()
        return rho

    def reset(self, st):
        """Restart from state st (x, y, u, v in user order) with cleared
        counters."""
        self.set_state(st)
        self.t = 0.0
        self.steps = 0
        self.plain_left = 0
        self.inst = True
        self.stats = {"surrogate": 0, "plain": 0, "checks": 0, "fallbacks": 0, "errors": []}

    def density_error(self, a, b):
        """Mean |rho_a - rho_b| / rho0 over the fluid."""
        d = (a[:self.nf] - b[:self.nf]).abs().mean() / self.info["rho0"]
        return float(d.item())

    def step(self):
        """One step of dt_factor x dt."""
        k = self.steps
        if self.plain_left > 0:
            self.set_state(self._plain_interval(self._instantaneous(self.state())))
            self.inst = True
            self.plain_left -= 1
            self.stats["plain"] += 1
        elif (k + 1) % self.check_every == 0:
            s0 = self.state()
            was_inst = self.inst
            self._surrogate_interval()
            s_sur = self.state()
            rho_sur = self.density()
            if not was_inst:
                s0 = {k2: s0[k2] for k2 in STATE}
                for c in ("u", "v"):
                    s0[c] = 0.5 * (s0[c] + s_sur[c])
            s_pl = self._plain_interval(s0)
            rho_pl = self.density(s_pl)
            err = self.density_error(rho_sur, rho_pl)
            self.stats["checks"] += 1
            self.stats["errors"].append(err)
            if not np.isfinite(err) or err > self.threshold:
                self.set_state(s_pl)
                self.inst = True
                self.stats["fallbacks"] += 1
                self.stats["plain"] += 1
                self.plain_left = self.check_every - 1
            else:
                self.set_state(s_sur)
                self.stats["surrogate"] += 1
        else:
            self._surrogate_interval()
            self.stats["surrogate"] += 1
        self.steps += 1
        self.t += self.dt * self.dt_factor

    def _instantaneous(self, st):
        """A state for the plain solver: states of the plain solver (and the
        fallback steps, which follow a plain state) already carry
        instantaneous velocities."""
        if self.inst:
            return st
        raise RuntimeError("surrogate: plain step from a leapfrog state outside a check")

    def run(self, nsteps):
        for _ in range(nsteps):
            self.step()
