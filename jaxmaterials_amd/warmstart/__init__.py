"""Learned warm start: network kernels (BF16 MFMA) and training on solver data."""

from .net import WarmStartNet, LAYERS, DEFAULT_NORM

__all__ = ["WarmStartNet", "LAYERS", "DEFAULT_NORM"]
