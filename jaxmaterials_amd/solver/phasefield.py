"""Phase-field fracture on the GPU, with the reference's function names and
arguments (src/jaxmaterials/solver/lippmann_schwinger.py:36-100, 334-547):

* ``compute_sigma_damaged(lmbda, mu, epsilon, d=None, k=1e-12)``
* ``compute_strain_energy(lmbda, mu, epsilon)``
* ``elasticity_solve(lmbda, mu, epsilon_bar, dvar, kk, grid_spec, ...)``
* ``phase_field_solve(HH, d_old, gc, lc, grid_spec, tolerance, maxiter)``

and ``staggered_steps``, the load-stepping loop of the reference's phase-field
script (tests/phasefield.py:190-225: phase field, then elasticity, then the
history update H = max(H, psi+), per macroscopic strain).  Fields are GPU
tensors: strains and Lame fields float32, damage and history float64 (the
reference runs the phase field in float64 and the elasticity in float32).
The tension/compression split uses the strain matrix as the reference forms
it (src/jaxmaterials/utilities.py:65-87; see jaxmaterials_amd/csrc/damage.hpp).
"""

import ctypes

import numpy as np
import torch

from .. import _lib
from .device import LSSolver

__all__ = ["compute_sigma_damaged", "compute_strain_energy", "elasticity_solve", "phase_field_solve",
           "staggered_steps"]


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _f32(t, shape, name):
    t = torch.as_tensor(t)
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a GPU tensor (no CPU path exists)")
    t = t.to(torch.float32).contiguous()
    if tuple(t.shape[-3:]) != tuple(shape[-3:]) or t.numel() != int(np.prod(shape)):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t.reshape(shape)


def _vox(shape):
    return np.array(shape[-3:], dtype=np.int32)


def compute_sigma_damaged(lmbda, mu, epsilon, d=None, k=1e-12):
    """Stress with the tension part degraded by (1 - d)^2 + k
    (lippmann_schwinger.py:36-74); epsilon [6, nx, ny, nz]."""
    shape = tuple(lmbda.shape[-3:])
    lam = _f32(lmbda, shape, "lmbda")
    m = _f32(mu, shape, "mu")
    eps = _f32(epsilon, (6,) + shape, "epsilon")
    dm = None if d is None else _f32(d, shape, "d")
    out = torch.empty_like(eps)
    vox = _vox(shape)
    torch.cuda.current_stream(eps.device).synchronize()
    rc = _lib.load().ls_damaged_stress(vox.ctypes.data_as(ctypes.c_void_p), _ptr(lam), _ptr(m), _ptr(eps), _ptr(dm),
                                       float(k), _ptr(out), None)
    if rc != 0:
        raise RuntimeError("ls_damaged_stress failed")
    return out


def compute_strain_energy(lmbda, mu, epsilon, history=None):
    """Tensile strain energy psi+ (lippmann_schwinger.py:77-100), float64
    [nx, ny, nz]; with ``history`` (float64 GPU tensor) also H = max(H, psi+)
    in place."""
    shape = tuple(lmbda.shape[-3:])
    lam = _f32(lmbda, shape, "lmbda")
    m = _f32(mu, shape, "mu")
    eps = _f32(epsilon, (6,) + shape, "epsilon")
    psi = torch.empty(shape, dtype=torch.float64, device=eps.device)
    if history is not None and (history.dtype != torch.float64 or not history.is_contiguous()
                                or tuple(history.shape) != shape):
        raise ValueError("history must be a contiguous float64 GPU tensor of the grid's shape")
    vox = _vox(shape)
    torch.cuda.current_stream(eps.device).synchronize()
    rc = _lib.load().ls_strain_energy(vox.ctypes.data_as(ctypes.c_void_p), _ptr(lam), _ptr(m), _ptr(eps), _ptr(psi),
                                      _ptr(history), None)
    if rc != 0:
        raise RuntimeError("ls_strain_energy failed")
    return psi


def elasticity_solve(lmbda, mu, epsilon_bar, dvar, kk, grid_spec, rtol=1e-6, atol=1e-20, depth=0, maxiter=32,
                     solver=None, graph=True):
    """Lippmann-Schwinger iteration with Anderson acceleration for the damaged
    law (lippmann_schwinger.py:334-462).  Returns (epsilon, sigma, iter) with
    GPU float32 [6, nx, ny, nz] fields; ``iter`` counts the updates and
    equals maxiter when the tolerance was not reached (no exception, as in the
    reference).  ``solver``: an LSSolver for this grid and depth to reuse;
    ``graph``: replay the iterations between host synchronisations as
    captured HIP graphs (fewer launches on small grids)."""
    shape = (grid_spec.nx, grid_spec.ny, grid_spec.nz)
    lam = _f32(lmbda, shape, "lmbda")
    m = _f32(mu, shape, "mu")
    d = None if dvar is None else _f32(dvar, shape, "dvar")
    eb = np.asarray(epsilon_bar, dtype=np.float64).reshape(6)
    if not np.any(eb):
        # no load: the reference's relative residual is 0 / 0, its loop does not
        # run and it returns the uniform start eps = 0 with its stress 0
        z = torch.zeros((6,) + shape, dtype=torch.float32, device=lam.device)
        return z, z.clone(), 0
    s = solver if solver is not None else LSSolver(grid_spec, nrhs=1, nmat=1, depth=depth)
    if s.depth != depth or s.nrhs != 1 or s.shape != shape:
        raise ValueError("solver does not match the grid, one load case and this depth")
    try:
        s.set_damage(d, kk)
        out = s.solve(lam, m, eb.astype(np.float32).reshape(1, 6), want_sigma=True, rtol=rtol,
                      atol=atol, maxiter=maxiter, check=False, graph=graph)
    finally:
        s.set_damage(None)
        if solver is None:
            s.close()
    return out["epsilon"][0], out["sigma"][0], int(out["iters"][0])


def phase_field_solve(HH, d_old, gc, lc, grid_spec, tolerance=1e-6, maxiter=1000, solver=None):
    """Fixed-point phase-field solve (lippmann_schwinger.py:465-547) in FP64.
    Returns (d, iter) with d a GPU float64 [nx, ny, nz] tensor."""
    shape = (grid_spec.nx, grid_spec.ny, grid_spec.nz)
    H = torch.as_tensor(HH).to(torch.float64).reshape(shape).contiguous()
    d0 = torch.as_tensor(d_old).to(torch.float64).reshape(shape).contiguous()
    if not H.is_cuda or not d0.is_cuda:
        raise RuntimeError("HH and d_old must be GPU tensors (no CPU path exists)")
    s = solver if solver is not None else LSSolver(grid_spec, nrhs=1, nmat=1, depth=0)
    try:
        d, it, _ = s.phase_field(H, d0, gc, lc, tolerance=tolerance, maxiter=maxiter)
    finally:
        if solver is None:
            s.close()
    return d, it


def staggered_steps(lmbda, mu, epsilon_bars, grid_spec, gc, lc, k_stab, depth=3, rtol=1e-6, atol=1e-20,
                    maxiter=1000, pf_tolerance=1e-9, pf_maxiter=10000, log=None):
    """The staggered load stepping of the reference's phase-field script
    (tests/phasefield.py:190-225): for each macroscopic strain, the phase field
    from the current history, the damaged elasticity solve, then
    H = max(H, psi+).  Returns a dict with the final damage and history, the
    final strain and stress, and per step the mean strain and stress and the
    iteration counts."""
    shape = (grid_spec.nx, grid_spec.ny, grid_spec.nz)
    lam = _f32(lmbda, shape, "lmbda")
    m = _f32(mu, shape, "mu")
    dev = lam.device
    d = torch.zeros(shape, dtype=torch.float64, device=dev)
    H = torch.zeros(shape, dtype=torch.float64, device=dev)
    s_el = LSSolver(grid_spec, nrhs=1, nmat=1, depth=depth)
    s_pf = s_el if depth == 0 else LSSolver(grid_spec, nrhs=1, nmat=1, depth=0)
    steps = []
    eps = sigma = None
    try:
        for i, eb in enumerate(epsilon_bars):
            d, it_pf = phase_field_solve(H, d, gc, lc, grid_spec, pf_tolerance, pf_maxiter, solver=s_pf)
            eps, sigma, it_el = elasticity_solve(lam, m, eb, d.to(torch.float32), k_stab, grid_spec, rtol=rtol,
                                                 atol=atol, depth=depth, maxiter=maxiter, solver=s_el)
            compute_strain_energy(lam, m, eps, history=H)
            rec = {"step": i, "pf_iterations": it_pf, "elasticity_iterations": it_el,
                   "eps_mean": eps.mean(dim=(1, 2, 3)).tolist(), "sigma_mean": sigma.mean(dim=(1, 2, 3)).tolist()}
            steps.append(rec)
            if log:
                log(rec)
    finally:
        s_el.close()
        if s_pf is not s_el:
            s_pf.close()
    return {"damage": d, "history": H, "epsilon": eps, "sigma": sigma, "steps": steps}
