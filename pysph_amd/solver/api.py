"""Public names of the solver layer (reference: solver/api.py)."""
from . import generators, utils  # noqa: F401
from .integrator import (EulerIntegrator, Integrator, LeapFrogIntegrator,  # noqa: F401
                         PredictorCorrectorIntegrator, RK2Integrator, RK4Integrator)
from .post_step_functions import PrintNeighborInformation  # noqa: F401
from .solver import Solver  # noqa: F401
from .sph_equation import SPHIntegration, SPHOperation  # noqa: F401
from .wcsph import wcsph_solver  # noqa: F401
from .wcsph_plan import WCSPHPlan  # noqa: F401
