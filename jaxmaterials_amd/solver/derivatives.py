"""Real-space divergence of a stress field on the GPU.

``relative_divergence(sigma, grid_spec)`` keeps the name, arguments and value
of the reference's real-space residual
(src/jaxmaterials/solver/lippmann_schwinger.py:103-115):
||D^- . sigma|| / (sqrt(N) ||<sigma>||_V) with the backward derivative of the
rotated staggered grid (src/jaxmaterials/solver/derivatives.py:8-96).
``backward_divergence`` returns the field D^- . sigma itself.  Both take a
CUDA float32 tensor of shape (6, nx, ny, nz) and run the HIP kernel of
``ls_relative_divergence``; they are an independent check of the solver's
Fourier residual, which equals the real-space one by Parseval.
"""

import ctypes

import numpy as np
import torch

from .. import _lib

__all__ = ["relative_divergence", "backward_divergence"]


def _run(sigma, grid_spec, want_div):
    if not (isinstance(sigma, torch.Tensor) and sigma.is_cuda and sigma.dtype == torch.float32):
        raise TypeError("sigma must be a CUDA float32 tensor")
    shape = (6, grid_spec.nx, grid_spec.ny, grid_spec.nz)
    if tuple(sigma.shape) != shape:
        raise ValueError(f"sigma has shape {tuple(sigma.shape)}, expected {shape}")
    sigma = sigma.contiguous()
    div = torch.empty((3,) + shape[1:], dtype=torch.float32, device=sigma.device) if want_div else None
    vox = np.array(shape[1:], dtype=np.int32)
    ext = np.array([grid_spec.Lx, grid_spec.Ly, grid_spec.Lz], dtype=np.float32)
    rel = ctypes.c_double(0.0)
    stream = torch.cuda.current_stream(sigma.device)
    rc = _lib.load().ls_relative_divergence(
        vox.ctypes.data_as(ctypes.c_void_p), ext.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(sigma.data_ptr()),
        ctypes.c_void_p(div.data_ptr() if div is not None else 0), ctypes.byref(rel),
        ctypes.c_void_p(stream.cuda_stream))
    if rc != 0:
        raise RuntimeError("ls_relative_divergence failed (see stderr)")
    return rel.value, div


def relative_divergence(sigma, grid_spec):
    """Relative divergence of sigma (lippmann_schwinger.py:103-115)."""
    return _run(sigma, grid_spec, False)[0]


def backward_divergence(sigma, grid_spec):
    """D^- . sigma, shape (3, nx, ny, nz) (derivatives.py:66-96)."""
    return _run(sigma, grid_spec, True)[1]
