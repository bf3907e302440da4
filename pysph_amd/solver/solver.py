"""Solver: operation list, integrator set-up and the time loop
(reference: solver/solver.py:26-574)."""
import logging
import os

import numpy as np

from ..base.kernels import CubicSplineKernel, DummyKernel
from ..sph import _ops
from ..sph.funcs import PositionStepping, XSPHCorrection
from .sph_equation import SPHIntegration

logger = logging.getLogger(__name__)


class Solver:
    def __init__(self, dim, integrator_type):
        self.particles = None
        self.integrator_type = integrator_type
        self.dim = dim
        self.default_kernel = CubicSplineKernel(dim=dim)
        self.initialize()
        self.setup_solver()
        self.with_cl = False

    def initialize(self):
        self.particles = None
        self.operation_dict = {}
        self.order = []
        self.t = 0
        self.count = 0
        self.execute_commands = None
        self.pre_step_functions = []
        self.post_step_functions = []
        self.pfreq = 100
        self.kernel_correction = -1
        self.pid = None
        self.eps = -1
        self.position_stepping_operations = {}
        self.print_properties = ["x", "u", "m", "h", "p", "rho"]
        if self.dim > 1:
            self.print_properties.extend(["y", "v"])
        if self.dim > 2:
            self.print_properties.extend(["z", "w"])
        self.fname = "solution"
        self.output_directory = "."
        self.detailed_output = False
        self.tf = 0.0
        self.dt = 0.0

    def switch_integrator(self, integrator_type):
        if self.particles is None:
            raise RuntimeError("There are no particles!")
        self.integrator_type = integrator_type
        self.setup_integrator(self.particles)

    def add_operation_step(self, types, xsph=False, eps=0.5):
        """PositionStepping on `types` (solver.py:124-146)."""
        updates = ["x", "y", "z"][:self.dim]
        self.add_operation(SPHIntegration(PositionStepping.withargs(dim=self.dim),
                                          on_types=types, updates=updates, id="step",
                                          kernel=DummyKernel(dim=self.dim)))

    def add_operation_xsph(self, eps, hks=False):
        """XSPHCorrection on the stepped types (solver.py:148-182)."""
        assert eps > 0, "Invalid value for XSPH epsilon: %f" % eps
        self.eps = eps
        assert "step" in self.operation_dict, "position stepping function does not exist!"
        types = self.operation_dict["step"].on_types
        updates = self.operation_dict["step"].updates
        self.add_operation(SPHIntegration(
            XSPHCorrection.withargs(eps=eps, hks=hks, dim=self.dim), from_types=types,
            on_types=types, updates=updates, id="xsph", kernel=self.default_kernel))

    def add_operation(self, operation, before=False, id=None):
        """Append, or insert before/after an existing id (solver.py:184-244)."""
        assert operation.id not in self.order, "Operation %s exists!" % operation.id
        assert operation.id not in self.operation_dict
        self.operation_dict[operation.id] = operation
        if id:
            msg = "The specified operation doesnt exist"
            assert id in self.operation_dict, msg + " in the calcs dict!"
            assert id in self.order, msg + " in the order list!"
            if before:
                self.order.insert(self.order.index(id), operation.id)
            else:
                self.order.insert(self.order.index(id) + 1, operation.id)
        else:
            self.order.append(operation.id)

    def replace_operation(self, id, operation):
        msg = "The specified operation doesnt exist"
        assert id in self.operation_dict, msg + " in the op dict!"
        assert id in self.order, msg + " in the order list!"
        self.operation_dict.pop(id)
        self.operation_dict[operation.id] = operation
        idx = self.order.index(id)
        self.order.insert(idx, operation.id)
        self.order.remove(id)

    def remove_operation(self, id_or_operation):
        id = id_or_operation if isinstance(id_or_operation, str) else id_or_operation.id
        assert id in self.operation_dict, "id doesnt exist!"
        assert id in self.order, "id doesnt exist!"
        self.order.remove(id)
        self.operation_dict.pop(id)

    def set_order(self, order):
        for equation_id in order:
            msg = "%s in order list does not exist!" % equation_id
            assert equation_id in self.order, msg
            assert equation_id in self.operation_dict, msg
        self.order = order

    def setup_position_step(self):
        pass

    def setup_integrator(self, particles=None):
        """Turn the operation order into calcs and set up the integrator
        (solver.py:320-366)."""
        if particles:
            self.particles = particles
            self.particles.kernel = self.default_kernel
            self.integrator = self.integrator_type(particles, calcs=[])
            for equation_id in self.order:
                operation = self.operation_dict[equation_id]
                if operation.kernel is None:
                    operation.kernel = self.default_kernel
                calcs = operation.get_calcs(particles, operation.kernel)
                self.integrator.calcs.extend(calcs)
            self.integrator.setup_integrator()

    def append_particle_arrrays(self, arrays):
        if not self.particles:
            print("Warning!, particles not defined")
            return
        for array in self.particles.arrays:
            for arr in arrays:
                if array.name == arr.name:
                    array.append_parray(arr)
        self.setup_integrator(self.particles)

    def set_final_time(self, tf):
        self.tf = tf

    def set_time_step(self, dt):
        self.dt = dt

    def set_print_freq(self, n):
        self.pfreq = n

    def set_output_fname(self, fname):
        self.fname = fname

    def set_output_printing_level(self, detailed_output):
        self.detailed_output = detailed_output

    def add_print_properties(self, props):
        for prop in props:
            if prop not in self.print_properties:
                self.print_properties.append(prop)

    def set_output_directory(self, path):
        self.output_directory = path

    def set_kernel_correction(self, kernel_correction):
        if kernel_correction != -1:
            raise NotImplementedError("kernel correction is not supported")
        self.kernel_correction = kernel_correction

    def set_cl(self, with_cl=False):
        if with_cl:
            raise NotImplementedError("the OpenCL path is replaced by the HIP path")

    def set_command_handler(self, callable, command_interval=1):
        self.execute_commands = callable
        self.command_interval = command_interval

    def solve(self, show_progress=False):
        """The time loop of solver.py:457-508."""
        self.count = 0
        while self.t < self.tf:
            self.t += self.dt
            self.count += 1
            self.particles.update()
            for func in self.pre_step_functions:
                func.eval(self, self.count)
            self.integrator.integrate(self.dt)
            for func in self.post_step_functions:
                func.eval(self, self.count)
            if self.pfreq and self.count % self.pfreq == 0:
                self.dump_output(*self.print_properties)
            if self.execute_commands is not None and self.count % self.command_interval == 0:
                self.execute_commands(self)
        # the jump check's device count of the last steps (cell.pyx:1422-1437)
        self.particles.cell_manager.verify_jumps()

    def dump_output(self, *print_properties):
        """One .npz per array: the print properties (or every property when
        detailed_output), dt, cell_size and np (solver.py:510-538)."""
        fname = self.fname + "_"
        cell_size = self.particles.cell_manager.cell_size
        paths = []
        for pa in self.particles.arrays:
            path = os.path.join(self.output_directory,
                                fname + pa.name + "_" + str(self.t) + ".npz")
            if self.detailed_output:
                props = pa.properties
                np.savez(path, dt=self.dt, **props)
            else:
                # get() returns the LocalReal particles, np counts them
                props = {p: pa.get(p) for p in print_properties if pa.has_array(p)}
                np.savez(path, dt=self.dt, cell_size=cell_size,
                         np=pa.num_real_particles, **props)
            paths.append(path)
        return paths

    def setup_solver(self):
        pass

    def conservation(self, gravity=(0.0, 0.0, 0.0), co=1.0, ro=1000.0, gamma=7.0):
        """FP64 mass, momentum and energy of the fluid arrays (no reference
        counterpart)."""
        plan = getattr(self.integrator, "plan", None)
        if plan is not None:
            prm = plan.prm
        else:
            prm = _ops.make_params(self.dim, co=co, ro=ro, gamma=gamma, g=gravity)
        fluids = [pa for pa in self.particles.arrays if pa.particle_type == 0]
        return _ops.conservation(prm, fluids)
