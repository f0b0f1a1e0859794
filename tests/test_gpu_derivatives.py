"""Real-space divergence kernels (ls_relative_divergence) against the FP64
oracle, and against the solver's Fourier residual (Parseval).

Mirrors test_relative_convergence (tests/test_lippmann_schwinger.py:50-61),
TestRelativeDivergence and TestFourierDivergence
(cuda/src/tests/test_derivatives.hh, cuda/src/tests/test_fourier.hh): the
real-space and Fourier relative divergences of the same stress agree.
"""

import numpy as np
import pytest
import torch

from jaxmaterials_amd.common import GridSpec
from jaxmaterials_amd.solver import backward_divergence, relative_divergence
from jaxmaterials_amd.solver.device import LSSolver
from oracle import ls_oracle as O

pytestmark = pytest.mark.gpu

GRIDS = {"even": ((64, 48, 32), (1.2, 0.8, 0.9)), "odd": ((57, 43, 37), (1.2, 0.8, 0.9)),
         "cube": ((32, 32, 32), (32.0, 32.0, 32.0))}


@pytest.mark.parametrize("grid", sorted(GRIDS))
def test_divergence_matches_oracle(gpu, grid):
    shape, ext = GRIDS[grid]
    gs = GridSpec(*shape, *ext)
    rng = np.random.default_rng(8741823)
    sigma = rng.normal(size=(6,) + shape).astype(np.float32) + np.array([1.0, 0.5, -0.3, 0.2, 0.1, -0.4],
                                                                         np.float32)[:, None, None, None]
    s_t = torch.from_numpy(sigma).to(gpu)
    div = backward_divergence(s_t, gs).cpu().numpy()
    rel = relative_divergence(s_t, gs)
    div_ref = O.backward_divergence(sigma.astype(np.float64), gs)
    rel_ref = O.relative_divergence(sigma.astype(np.float64), gs)
    err = np.abs(div - div_ref).max() / np.abs(div_ref).max()
    print(f"divergence {grid}: max rel err {err:.2e}, relative divergence {rel:.6e} vs {rel_ref:.6e}", flush=True)
    assert err < 1e-5
    assert abs(rel - rel_ref) <= 1e-5 * rel_ref


@pytest.mark.parametrize("grid", ["even", "odd"])
def test_real_space_residual_equals_solver_residual(gpu, grid):
    """At convergence the solver's Fourier residual R of the returned strain
    (rel_ref = 1 makes the reported residual R itself) equals the real-space
    relative divergence of the returned stress."""
    shape, ext = GRIDS[grid]
    gs = GridSpec(*shape, *ext)
    rng = np.random.default_rng(784173)
    lam = torch.from_numpy(rng.uniform(0.5, 1.5, size=shape).astype(np.float32)).to(gpu)
    mu = torch.from_numpy(rng.uniform(0.5, 1.5, size=shape).astype(np.float32)).to(gpu)
    s = LSSolver(gs, nrhs=1)
    eb = np.array([[1.0, 0.4, 0.3, 0.1, -0.4, 0.7]], np.float32)
    out = s.solve(lam, mu, eb, rtol=1e-3, atol=0.0, maxiter=100, rel_ref=1.0, want_sigma=True)
    R_fourier = float(out["residual"][0])
    R_real = relative_divergence(out["sigma"][0], gs)
    print(f"residual {grid}: Fourier {R_fourier:.6e} real space {R_real:.6e}", flush=True)
    assert abs(R_fourier - R_real) <= 2e-3 * R_real
    s.close()
