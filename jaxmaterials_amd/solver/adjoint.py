"""Differentiable Lippmann-Schwinger solve: a torch.autograd.Function whose
backward pass is a hand-written adjoint (a second LS solve on the GPU).

The reference differentiates only forward-mode, only with respect to the
macroscopic strain, through ``jax.lax.while_loop`` (src/driver.py:73-76,
tests/test_yc.py:135-137); its adjoint derivation is unfinished
(doc/linear_elasticity.tex:404-447).  Here the derivative with respect to the
per-voxel Lame fields is the adjoint of the discrete problem solved by the
iteration:

    eps = eps_bar + e,  e compatible with zero mean,  Gamma0 (C eps) = 0.

For a loss L with dL = sum_x G(x) . d eps(x) (plain partials G), write the
Voigt-weighted gradient g = G / w, w = (1, 1, 1, 2, 2, 2).  Because Gamma0 is
symmetric in the weighted product and C is symmetric, the adjoint state q
solves the same kind of problem: q compatible with zero mean and
Gamma0 (C q - g) = 0, i.e. the LS iteration with polarisation tau = -g and
eps_bar = 0.  Then

    dL/dlambda(x) = -tr q(x) tr eps(x)
    dL/dmu(x)     = -2 <q(x), eps(x)>_w
    dL/deps_bar_j = w_j sum_x (g - C q)_j(x).

The loss may also depend on lambda, mu directly (e.g. through sigma = C eps);
torch's autograd adds that part, so ``C_eff`` losses are written in torch on
top of ``solve_strain``.
"""

import numpy as np
import torch

from .device import LSSolver

__all__ = ["LSSolve", "solve_strain", "effective_tensor_torch", "hooke_torch", "VOIGT_W"]

VOIGT_W = (1.0, 1.0, 1.0, 2.0, 2.0, 2.0)


def hooke_torch(lmbda, mu, eps):
    """sigma = 2 mu eps + lambda tr(eps) [i<3] (lippmann_schwinger.py:20-33), differentiable."""
    tr = eps[:, 0] + eps[:, 1] + eps[:, 2]
    sig = 2.0 * mu * eps
    lam_tr = (lmbda * tr).unsqueeze(1)
    return torch.cat([sig[:, :3] + lam_tr, sig[:, 3:]], dim=1)


class LSSolve(torch.autograd.Function):
    """eps = LS(lambda, mu, eps_bar) for a batch of load cases on one grid."""

    @staticmethod
    def forward(ctx, lmbda, mu, eps_bar, solver, rtol, maxiter, adjoint_rtol):
        eb = eps_bar.detach().to(torch.float32).cpu().numpy()
        out = solver.solve(lmbda.detach().contiguous(), mu.detach().contiguous(), eb, rtol=rtol, maxiter=maxiter)
        eps = out["epsilon"]
        ctx.solver = solver
        ctx.maxiter = maxiter
        ctx.adjoint_rtol = adjoint_rtol
        # the eps_bar gradient is returned on eps_bar's own device and dtype
        ctx.eb_device, ctx.eb_dtype = eps_bar.device, eps_bar.dtype
        ctx.save_for_backward(lmbda, mu, eps)
        return eps

    @staticmethod
    def backward(ctx, G):
        lmbda, mu, eps = ctx.saved_tensors
        solver = ctx.solver
        R = eps.shape[0]
        w = torch.tensor(VOIGT_W, device=eps.device, dtype=torch.float32).view(1, 6, 1, 1, 1)
        g = (G / w).contiguous()
        tau = (-g).contiguous()
        out = solver.solve(lmbda.detach().contiguous(), mu.detach().contiguous(), np.zeros((R, 6), np.float32),
                           tau=tau, rtol=ctx.adjoint_rtol, maxiter=ctx.maxiter, absolute_residual=True)
        q = out["epsilon"]
        tr_q = q[:, 0] + q[:, 1] + q[:, 2]
        tr_e = eps[:, 0] + eps[:, 1] + eps[:, 2]
        d_lam = -(tr_q * tr_e).sum(0)
        d_mu = -2.0 * (q * eps * w).sum(dim=(0, 1))
        Cq = hooke_torch(lmbda.unsqueeze(0), mu.unsqueeze(0), q)
        d_eb = ((g.double() - Cq.double()).sum(dim=(2, 3, 4)) * w.view(1, 6).double())
        d_eb = d_eb.to(device=ctx.eb_device, dtype=ctx.eb_dtype)
        return d_lam.reshape(lmbda.shape), d_mu.reshape(mu.shape), d_eb, None, None, None, None


def solve_strain(lmbda, mu, eps_bar, solver=None, rtol=1e-6, maxiter=2000, adjoint_rtol=None):
    """Differentiable strain fields [R, 6, nx, ny, nz] of the load cases eps_bar [R, 6].

    lmbda, mu: GPU float32 [nx, ny, nz] (requires_grad allowed); eps_bar: tensor [R, 6].
    """
    if solver is None:
        raise ValueError("pass an LSSolver built for this grid and number of load cases")
    if not isinstance(eps_bar, torch.Tensor):
        eps_bar = torch.as_tensor(np.asarray(eps_bar), dtype=torch.float64)
    return LSSolve.apply(lmbda, mu, eps_bar, solver, rtol, maxiter, adjoint_rtol if adjoint_rtol else rtol)


def effective_tensor_torch(lmbda, mu, solver, rtol=1e-6, maxiter=2000, adjoint_rtol=None):
    """Differentiable 6x6 C_eff (row k = <sigma> under eps_bar = e_k, data.py:178-186)."""
    eb = torch.eye(6, dtype=torch.float64)
    eps = solve_strain(lmbda, mu, eb, solver=solver, rtol=rtol, maxiter=maxiter, adjoint_rtol=adjoint_rtol)
    sig = hooke_torch(lmbda.unsqueeze(0), mu.unsqueeze(0), eps)
    return sig.to(torch.float64).mean(dim=(2, 3, 4))


def make_solver(grid_spec, nrhs=6, depth=0):
    """LSSolver for solve_strain / effective_tensor_torch."""
    return LSSolver(grid_spec, nrhs=nrhs, depth=depth)
