"""Seeded cross-ply fibre microstructures (RSA placement + GPU rasteriser).

Wraps ``ls_crossply_place`` / ``ls_crossply_lame`` of liblippmannschwinger.so;
the placement rules are documented in jaxmaterials_amd/csrc/microstructure.hip.
The rasterisation follows the reference's material_properties
(src/jaxmaterials/data/data.py:267-314).
"""

import ctypes

import numpy as np
import torch

from .. import _lib

__all__ = ["crossply_place", "crossply_lame", "lame_from_E_nu", "CONFIG_MODULI"]

# Phase moduli of the cross-ply benchmark configurations:
# matrix E = 3, nu = 0.35 -> (lambda, mu) = (2.593, 1.111);
# fibre E = 70, nu = 0.2 -> (lambda, mu) = (19.44, 29.17).
CONFIG_MODULI = {"lam_matrix": 2.593, "mu_matrix": 1.111, "lam_fibre": 19.44, "mu_fibre": 29.17}


def lame_from_E_nu(E, nu):
    """(lambda, mu) of an isotropic material."""
    mu = E / (2.0 * (1.0 + nu))
    return 2.0 * mu * nu / (1.0 - 2.0 * nu), mu


def _arrs(shape, extents):
    return np.array(shape, dtype=np.int32), np.array(extents, dtype=np.float32)


def crossply_place(shape, extents, nlayers, radius, vf, seed):
    """RSA fibre centres: float64 [n, 3] rows (layer, in-plane coordinate, z)."""
    lib = _lib.load()
    vox, ext = _arrs(shape, extents)
    cap = 1 << 16
    out = np.zeros((cap, 3), dtype=np.float64)
    n = lib.ls_crossply_place(vox.ctypes.data_as(ctypes.c_void_p), ext.ctypes.data_as(ctypes.c_void_p),
                              int(nlayers), float(radius), float(vf), ctypes.c_ulonglong(seed),
                              out.ctypes.data_as(ctypes.c_void_p), cap)
    if n < 0:
        raise RuntimeError("ls_crossply_place failed")
    return out[: min(n, cap)].copy()


def crossply_lame(shape, extents, nlayers, radius, vf, seed, lam_matrix=2.593, mu_matrix=1.111,
                  lam_fibre=19.44, mu_fibre=29.17, device="cuda"):
    """Lame fields of a seeded cross-ply microstructure, rasterised on the GPU.

    Returns (lmbda, mu, achieved_vf, n_fibres) with lmbda, mu float32 GPU tensors
    of the given shape.
    """
    lib = _lib.load()
    vox, ext = _arrs(shape, extents)
    lam = torch.empty(tuple(shape), dtype=torch.float32, device=device)
    mu = torch.empty(tuple(shape), dtype=torch.float32, device=device)
    if not lam.is_cuda:
        raise RuntimeError("crossply_lame rasterises on the GPU; device must be a GPU")
    vf_out = ctypes.c_double(0.0)
    torch.cuda.synchronize()
    n = lib.ls_crossply_lame(vox.ctypes.data_as(ctypes.c_void_p), ext.ctypes.data_as(ctypes.c_void_p),
                             int(nlayers), float(radius), float(vf), ctypes.c_ulonglong(seed),
                             float(lam_matrix), float(mu_matrix), float(lam_fibre), float(mu_fibre),
                             ctypes.c_void_p(lam.data_ptr()), ctypes.c_void_p(mu.data_ptr()),
                             ctypes.byref(vf_out))
    if n < 0:
        raise RuntimeError("ls_crossply_lame failed (see stderr)")
    torch.cuda.synchronize()
    return lam, mu, vf_out.value, n
