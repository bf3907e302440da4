"""pysph_amd: PySPH's particle step (cell-binned neighbour search, WCSPH
pair sums, Euler/RK2 integration) on AMD MI355X (gfx950).

Subpackages mirror the reference's layout: `base` (ParticleArray,
Particles, CellManager, NNPSManager, kernels), `sph` (SPH functions,
SPHCalc), `solver` (operations, integrators, Solver).  All compute runs in
libsph_mi355x.so (csrc/, C ABI in include/sph_mi355x.h).
"""
__version__ = "0.1.0"
