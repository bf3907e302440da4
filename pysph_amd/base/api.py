"""Public names of the base layer (reference: base/api.py)."""
from .carray import DoubleArray, FloatArray, IntArray, LongArray  # noqa: F401
from .cell import (CellManager, PeriodicDomain, construct_immediate_neighbor_list,  # noqa: F401
                   find_cell_id, py_find_cell_id, py_real_to_int, real_to_int)
from .kernels import CubicSplineKernel, DummyKernel, KernelBase  # noqa: F401
from .nnps import (FixedDestNbrParticleLocator, NbrParticleLocatorBase,  # noqa: F401
                   NeighborLocatorType, NNPSManager, VarHNbrParticleLocator, neighbor_csr)
from .particle_array import ParticleArray, get_particle_array  # noqa: F401
from .particle_types import (GhostParticle, LocalDummy, LocalReal, ParticleType,  # noqa: F401
                             RemoteDummy, RemoteReal)
from .particles import Particles  # noqa: F401
from .point import IntPoint, Point  # noqa: F401

Fluid = ParticleType.Fluid
Solid = ParticleType.Solid
Probe = ParticleType.Probe
