"""Edge-MLP message-passing surrogate of the WCSPH step (larger time steps,
checked against and falling back to the plain step)."""
from .model import EdgeMLP  # noqa: F401
from .graph import ParticleGraph  # noqa: F401
from .solver import SurrogateWCSPH  # noqa: F401
