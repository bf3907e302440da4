"""Integrators (reference: solver/integrator.py:9-645).

Two execution paths give the same arithmetic:

* the generic path runs every calc through SPHCalc.sph and performs the
  reference's array bookkeeping (`_p_0` shared initial arrays, `_k{n}_...`
  stage arrays, cumulative per-calc updates) with GPU update kernels;
* the fused path (wcsph_plan.WCSPHPlan) is chosen automatically when the
  calcs form the standard WCSPH set (SPHRho, TaitEquation,
  MomentumEquation, [GravityForce], PositionStepping, [XSPHCorrection]) with
  one moving array.  It evaluates density+EOS in one pass per array and
  momentum+XSPH+the stage update in one pass, keeping the same per-calc
  update order.

Set `integrator.fused = False` before setup_integrator() to force the
generic path.
"""
import logging

import torch

from ..sph import _ops

logger = logging.getLogger(__name__)


class Integrator:
    def __init__(self, particles=None, calcs=(), pcalcs=()):
        self.particles = particles
        self.calcs = list(calcs)
        self.pcalcs = list(pcalcs)
        self.nsteps = 1
        self.cstep = 1
        self.setup_done = False
        self.rupdate_list = []
        self.fused = True
        self.plan = None
        self.initial_props = {}
        self.k_props = {}

    def set_rupdate_list(self):
        self.rupdate_list = [[] for _ in self.particles.arrays]

    def setup_integrator(self):
        """Create `_p_0` and `_k{n}_{p}{i}{j}` properties (integrator.py:
        229-337) and pick the execution path."""
        self.arrays = self.particles.arrays
        self.initial_props = {}
        self.k_props = {}
        calcs = self.calcs
        self.pcalcs = [c for c in calcs if c.tag == "position"]
        self.ncalcs = [c for c in calcs if c.tag != "position"]
        self.icalcs = [c for c in calcs if c.tag != "position" and c.integrates]
        for calc in calcs:
            calc.initial_props = []
            calc.dst_writes = {}
        for i, calc in enumerate(calcs):
            pa = self.arrays[calc.dnum]
            self.initial_props[calc.id] = []
            self.k_props[calc.id] = {}
            for j, prop in enumerate(calc.updates):
                prop_initial = "_" + prop + "_0"
                if calc.integrates:
                    pa.add_property({"name": prop_initial})
                    calc.initial_props.append(prop_initial)
                self.initial_props[calc.id].append(prop_initial)
                for l in range(self.nsteps):
                    k_num = "k" + str(l + 1)
                    if calc.integrates:
                        self.k_props[calc.id].setdefault(k_num, [])
                        k_name = "_" + k_num + "_" + prop + str(i) + str(j)
                        pa.add_property({"name": k_name})
                        calc.dst_writes.setdefault(k_num, []).append(k_name)
                        self.k_props[calc.id][k_num].append(k_name)
        self.set_rupdate_list()
        self.plan = None
        if self.fused:
            from .wcsph_plan import WCSPHPlan
            self.plan = WCSPHPlan.match(self.particles, calcs)
            if self.plan is not None:
                logger.info("integrator: fused WCSPH plan on array %s", self.plan.fluid.name)
        # the fused plan keeps its own stage buffers: the _p_0 / _k arrays
        # stay allocated (as the reference creates them) but are not carried
        # through re-binnings, and hold no meaningful values while it runs
        scratch = {}
        for calc in calcs:
            pa = self.arrays[calc.dnum]
            names = scratch.setdefault(id(pa), set())
            names.update(calc.initial_props)
            for ks in self.k_props[calc.id].values():
                names.update(ks)
        for pa in self.arrays:
            pa.__dict__["_unpermuted"] = set(scratch.get(id(pa), ())) if self.plan else set()
        self.setup_done = True

    # ------------------------------------------------------------ generic path
    def _apply(self, pa, mode, s, ops):
        _ops.update(pa.dtype, pa.get_number_of_particles(), mode, s, ops)

    def set_initial_arrays(self):
        """`_p_0 = p` for every integrating calc (integrator.py:339-367)."""
        for calc in self.calcs:
            if not calc.integrates:
                continue
            pa = self.arrays[calc.dnum]
            ops = [(pa.get_carray(self.initial_props[calc.id][j]), pa.get_carray(prop), None, None)
                   for j, prop in enumerate(calc.updates)]
            self._apply(pa, 3, 0.0, ops)

    def reset_current_arrays(self, calcs):
        """`p = _p_0` (integrator.py:369-392)."""
        for calc in calcs:
            if not calc.integrates:
                continue
            pa = self.arrays[calc.dnum]
            ops = [(pa.get_carray(prop), pa.get_carray(self.initial_props[calc.id][j]), None, None)
                   for j, prop in enumerate(calc.updates)]
            self._apply(pa, 3, 0.0, ops)

    def eval(self, calcs):
        """Evaluate each calc in order: integrating calcs write their k
        arrays, the others assign their updates at once (integrator.py:
        432-470)."""
        particles = self.particles
        k_num = "k" + str(self.cstep)
        for calc in calcs:
            if calc.integrates:
                calc.sph(*calc.dst_writes[k_num])
            else:
                calc.sph(*calc.updates)
                particles.barrier()
                self.rupdate_list[calc.dnum] = [calc.updates]
                particles.update_remote_particle_properties(self.rupdate_list)
        particles.barrier()

    def step(self, calcs, dt):
        """`p = p + k*dt` per integrating calc, in order (integrator.py:
        472-514)."""
        k_num = "k" + str(self.cstep)
        for calc in calcs:
            if not calc.integrates:
                continue
            pa = self.arrays[calc.dnum]
            ops = []
            for j, prop in enumerate(calc.updates):
                cur = pa.get_carray(prop)
                ops.append((cur, cur, pa.get_carray(self.k_props[calc.id][k_num][j]), None))
            self._apply(pa, 0, dt, ops)
            if any(p in ("x", "y", "z") for p in calc.updates):
                pa.set_dirty(True)
        self.cstep += 1

    def do_step(self, calcs, dt):
        self.eval(calcs)
        self.particles.barrier()
        self.reset_current_arrays(calcs)
        self.step(calcs, dt)
        self.particles.barrier()

    def integrate(self, dt):
        raise NotImplementedError

    def _moved(self):
        for calc in self.calcs:
            if calc.integrates and any(p in ("x", "y", "z") for p in calc.updates):
                self.arrays[calc.dnum].set_dirty(True)


class EulerIntegrator(Integrator):
    """X(t+h) = X + h F(X) (integrator.py:525-573)."""

    def __init__(self, particles=None, calcs=()):
        Integrator.__init__(self, particles, calcs)
        self.nsteps = 1

    def final_step(self, calc, dt):
        """`p = p_0 + k1*dt; p_0 = p` (integrator.py:534-551)."""
        pa = self.arrays[calc.dnum]
        ops = []
        for i, prop in enumerate(calc.updates):
            ops.append((pa.get_carray(prop), pa.get_carray(self.initial_props[calc.id][i]),
                        pa.get_carray(self.k_props[calc.id]["k1"][i]), None))
        self._apply(pa, 2, dt, ops)

    def integrate(self, dt):
        if self.plan is not None:
            self.plan.step(dt, euler=True)
            return
        self.set_initial_arrays()
        self.eval(self.calcs)
        for calc in self.calcs:
            if calc.integrates:
                self.final_step(calc, dt)
        self._moved()
        self.particles.update()


class RK2Integrator(Integrator):
    """Heun: K1 = F(X), K2 = F(X + h K1), X += h/2 (K1 + K2)
    (integrator.py:581-645)."""

    def __init__(self, particles=None, calcs=()):
        Integrator.__init__(self, particles, calcs)
        self.nsteps = 2

    def final_step(self, calc, dt):
        """`p = p_0 + 0.5*dt*(k1 + k2); p_0 = p` (integrator.py:597-619)."""
        pa = self.arrays[calc.dnum]
        ops = []
        for j, prop in enumerate(calc.updates):
            ops.append((pa.get_carray(prop), pa.get_carray(self.initial_props[calc.id][j]),
                        pa.get_carray(self.k_props[calc.id]["k1"][j]),
                        pa.get_carray(self.k_props[calc.id]["k2"][j])))
        self._apply(pa, 1, 0.5 * dt, ops)

    def integrate(self, dt):
        if self.plan is not None:
            self.plan.step(dt)
            return
        self.set_initial_arrays()
        self.do_step(self.calcs, dt)
        self.particles.update()
        self.eval(self.calcs)
        for calc in self.calcs:
            if calc.integrates:
                self.final_step(calc, dt)
        self.cstep = 1
        self._moved()
        self.particles.update()


class RK4Integrator(Integrator):
    """Classical RK4: K1 = F(X), K2 = F(X + h/2 K1), K3 = F(X + h/2 K2),
    K4 = F(X + h K3), X += h/6 (K1 + 2 K2 + 2 K3 + K4)
    (integrator.py:650-724)."""

    def __init__(self, particles=None, calcs=()):
        Integrator.__init__(self, particles, calcs)
        self.nsteps = 4
        self.fused = False

    def final_step(self, calc, dt):
        """`p = p_0 + dt/6*(k1 + 2*k2 + 2*k3 + k4); p_0 = p`
        (integrator.py:669-689), summed in the reference's order."""
        pa = self.arrays[calc.dnum]
        kp = self.k_props[calc.id]
        tmp = []
        partial = []
        for j, prop in enumerate(calc.updates):
            t = torch.empty_like(pa.get_carray(prop))
            tmp.append(t)
            partial.append((t, pa.get_carray(kp["k1"][j]), pa.get_carray(kp["k2"][j]),
                            pa.get_carray(kp["k3"][j])))
        self._apply(pa, 5, 0.0, partial)
        ops = []
        for j, prop in enumerate(calc.updates):
            ops.append((pa.get_carray(prop), pa.get_carray(self.initial_props[calc.id][j]),
                        tmp[j], pa.get_carray(kp["k4"][j])))
        self._apply(pa, 1, dt / 6.0, ops)

    def integrate(self, dt):
        self.set_initial_arrays()
        self.do_step(self.calcs, 0.5 * dt)
        self._moved()
        self.particles.update()
        self.do_step(self.calcs, 0.5 * dt)
        self._moved()
        self.particles.update()
        self.do_step(self.calcs, dt)
        self._moved()
        self.particles.update()
        self.eval(self.calcs)
        for calc in self.calcs:
            if calc.integrates:
                self.final_step(calc, dt)
        self.cstep = 1
        self._moved()
        self.particles.update()


class PredictorCorrectorIntegrator(Integrator):
    """Predict X(t + h/2) = X + h/2 F(X), correct X(t + h/2) = X + h/2
    F(X(t + h/2)), then X(t + h) = 2 X(t + h/2) - X (integrator.py:726-798)."""

    def __init__(self, particles=None, calcs=()):
        Integrator.__init__(self, particles, calcs)
        self.nsteps = 2
        self.fused = False

    def final_step(self, calc, dt):
        """`p = 2*p - p_0; p_0 = p` (integrator.py:756-772)."""
        pa = self.arrays[calc.dnum]
        ops, copies = [], []
        for j, prop in enumerate(calc.updates):
            cur = pa.get_carray(prop)
            init = pa.get_carray(self.initial_props[calc.id][j])
            ops.append((cur, cur, init, None))
            copies.append((init, cur, None, None))
        self._apply(pa, 7, 0.0, ops)
        self._apply(pa, 3, 0.0, copies)

    def integrate(self, dt):
        self.set_initial_arrays()
        self.do_step(self.calcs, 0.5 * dt)
        self._moved()
        self.particles.update()
        self.do_step(self.calcs, 0.5 * dt)
        self._moved()
        self.particles.update()
        for calc in self.calcs:
            if calc.integrates:
                self.final_step(calc, dt)
        self._moved()
        self.particles.update()
        self.cstep = 1


class LeapFrogIntegrator(Integrator):
    """Leap frog (integrator.py:800-912): step the non-position calcs, then
    the position calcs with the new velocities, add 0.5 h^2 F_0 to the
    positions, re-evaluate and correct v += 0.5 h (F - F_0)."""

    def __init__(self, particles=None, calcs=()):
        Integrator.__init__(self, particles, calcs)
        self.nsteps = 2
        self.fused = False

    def add_correction_for_position(self, dt):
        """positions += 0.5*dt*dt * k1 of each velocity calc
        (integrator.py:829-858)."""
        if not self.pcalcs:
            return
        pos_calc = self.pcalcs[0]
        ppa = self.arrays[pos_calc.dnum]
        for calc in self.icalcs:
            if calc.tag != "velocity":
                continue
            pa = self.particles.arrays[calc.dnum]
            ops = []
            for j in range(len(calc.updates)):
                cur = ppa.get_carray(pos_calc.updates[j])
                ops.append((cur, cur, pa.get_carray(self.k_props[calc.id]["k1"][j]), None))
            self._apply(ppa, 0, 0.5 * dt * dt, ops)
        ppa.set_dirty(True)

    def final_step(self, calc, dt):
        """`p = p + 0.5*dt*(k2 - k1)` (integrator.py:860-883)."""
        pa = self.arrays[calc.dnum]
        ops = []
        for j, prop in enumerate(calc.updates):
            cur = pa.get_carray(prop)
            ops.append((cur, cur, pa.get_carray(self.k_props[calc.id]["k2"][j]),
                        pa.get_carray(self.k_props[calc.id]["k1"][j])))
        self._apply(pa, 8, 0.5 * dt, ops)

    def integrate(self, dt):
        self.set_initial_arrays()
        self.do_step(self.ncalcs, dt)
        self.cstep = 1
        self.do_step(self.pcalcs, dt)
        self.add_correction_for_position(dt)
        self._moved()
        self.particles.barrier()
        self.particles.update()
        self.eval(self.ncalcs)
        for calc in self.icalcs:
            self.final_step(calc, dt)
        self.cstep = 1
