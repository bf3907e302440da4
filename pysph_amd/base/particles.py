"""Particles: the arrays of a simulation with their cell and neighbour
managers (reference: base/particles.py:23-280).

MPI domain decomposition (in_parallel) is not part of this single-GPU
build; asking for it raises NotImplementedError.  variable_h=True gives the
symmetric variable-h neighbour sets (VarHNbrParticleLocator) to the locators
and to every SPH operation on these arrays.
"""
from .cell import CellManager
from .nnps import NeighborLocatorType, NNPSManager
from .particle_array import ParticleArray, get_particle_array  # noqa: F401
from .particle_types import ParticleType

SPHNeighborLocator = NeighborLocatorType.SPHNeighborLocator


class Particles:
    def __init__(self, arrays=(), in_parallel=False, variable_h=False, load_balancing=True,
                 update_particles=True, locator_type=SPHNeighborLocator,
                 periodic_domain=None, min_cell_size=-1):
        if in_parallel:
            raise NotImplementedError("in_parallel (MPI) runs are not supported; one GPU "
                                      "holds the whole particle set")
        assert len(arrays) > 0, "Particles must be given some arrays!"
        prec = arrays[0].cl_precision
        for arr in arrays[1:]:
            assert arr.cl_precision == prec, "All arrays must have the same cl_precision"
        self.variable_h = variable_h
        self.in_parallel = in_parallel
        self.load_balancing = load_balancing
        self.locator_type = locator_type
        self.arrays = list(arrays)
        self.kernel = None
        self.cell_manager = CellManager(arrays_to_bin=self.arrays, min_cell_size=min_cell_size,
                                        periodic_domain=periodic_domain)
        self.nnps_manager = NNPSManager(cell_manager=self.cell_manager, variable_h=variable_h,
                                        locator_type=locator_type)
        self.correction_manager = None
        self.misc_prop_update_functions = []
        if update_particles:
            self.update()

    def update(self, cache_neighbors=False):
        """Mark locators and the cell manager dirty when arrays moved; the
        re-binning itself happens lazily before the next neighbour pass, as
        the reference's cache rebuild does (particles.py:159-200)."""
        err = self.nnps_manager.py_update()
        assert err != -1, "NNPSManager update failed! "
        self.evaluate_misc_properties()
        if self.correction_manager:
            self.correction_manager.update()

    def evaluate_misc_properties(self):
        for func in self.misc_prop_update_functions:
            func.eval()

    def add_misc_function(self, func, operation, kernel):
        calcs = operation.get_calcs(self, kernel)
        self.misc_prop_update_functions.append(func(calcs))

    def get_named_particle_array(self, name):
        for array in self.arrays:
            if array.name == name:
                return array
        print("Array %s does not exist!" % name)

    def update_remote_particle_properties(self, props=None):
        """No remote particles on one GPU."""

    def barrier(self):
        """Single process: nothing to synchronise."""

    @classmethod
    def get_neighbor_particle_locator(cls, src, dst, locator_type=SPHNeighborLocator,
                                      variable_h=False, radius_scale=2.0):
        cell_manager = CellManager(arrays_to_bin=[src] if src is dst else [src, dst])
        nnps_manager = NNPSManager(cell_manager, locator_type=locator_type,
                                   variable_h=variable_h)
        return nnps_manager.get_neighbor_particle_locator(src, dst, radius_scale)

    def get_cl_precision(self):
        return self.arrays[0].cl_precision
