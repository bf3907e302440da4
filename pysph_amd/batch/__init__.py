"""Batched multi-scene WCSPH engine."""
from .engine import BatchedWCSPH, Scene, scene_from_arrays  # noqa: F401
