"""Solver entry points."""

from .lippmann_schwinger import lippmann_schwinger_cuda, effective_tensor  # noqa: F401
from .derivatives import relative_divergence, backward_divergence  # noqa: F401

__all__ = ["lippmann_schwinger_cuda", "effective_tensor", "relative_divergence", "backward_divergence"]
