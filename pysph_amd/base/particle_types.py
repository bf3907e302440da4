"""Particle types (reference: base/particle_types.py:1-23) and particle tags
(base/particle_array.pxd:20-26)."""


class ParticleType:
    """Enumeration of particle array types; not instantiable."""
    Fluid = 0
    Solid = 1
    DummyFluid = 2
    Probe = 3

    def __init__(self):
        raise SystemError("Do not instantiate the ParticleType class")


# ParticleTag values (particle_array.pxd:20-26)
LocalReal = 0
LocalDummy = 1
RemoteReal = 2
RemoteDummy = 3
GhostParticle = 4
