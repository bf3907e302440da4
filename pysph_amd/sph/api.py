"""Public names of the SPH layer (reference: sph/api.py)."""
from .funcs import (GravityForce, MomentumEquation, MoveCircleX, MoveCircleY,  # noqa: F401
                    PositionStepping, SPHDensityRate, SPHRho, TaitEquation, XSPHCorrection)
from .sph_calc import SPHCalc  # noqa: F401
from .sph_func import (CSPHFunctionParticle, Function, SPHFunction,  # noqa: F401
                       SPHFunctionParticle)
