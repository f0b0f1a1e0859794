"""The no-allocation entry point ls_solve_noalloc (buffers in the library's
per-GPU scratch array) against the allocating path it is a variant of
(ls_solver_create + ls_solver_solve through LSSolver): identical iterations,
strain and FP64 volume averages, on two calls in a row with the second on a
smaller problem, and a clear refusal of a problem larger than the array."""

import ctypes

import numpy as np
import pytest
import torch

from jaxmaterials_amd import _lib
from jaxmaterials_amd.common import GridSpec
from jaxmaterials_amd.data import crossply_lame
from jaxmaterials_amd.solver.device import LSSolver

pytestmark = pytest.mark.gpu


def _p(a):
    return ctypes.c_void_p(a.data_ptr()) if isinstance(a, torch.Tensor) else a.ctypes.data_as(ctypes.c_void_p)


def _noalloc(shape, lam, mu, eb, depth, rtol=1e-6, maxiter=2000, want_sigma=True):
    lib = _lib.load()
    R = eb.shape[0]
    vox = np.array(shape, np.int32)
    ext = np.array([float(n) for n in shape], np.float32)
    eps = torch.full((R, 6) + shape, float("nan"), dtype=torch.float32, device=lam.device)
    sig = torch.full_like(eps, float("nan")) if want_sigma else None
    iters = np.zeros(R, np.int32)
    res = np.zeros(R, np.float64)
    smean = np.zeros((R, 6), np.float64)
    torch.cuda.synchronize()
    rc = lib.ls_solve_noalloc(_p(vox), _p(ext), R, depth, _p(lam), _p(mu), _p(np.ascontiguousarray(eb, np.float32)),
                              _p(eps), _p(sig) if sig is not None else None, rtol, 0.0, maxiter, _p(iters), _p(res),
                              _p(smean), None)
    return rc, eps, sig, iters, res, smean


def _reference(shape, lam, mu, eb, depth, rtol=1e-6, maxiter=2000):
    s = LSSolver(GridSpec(*shape, *[float(n) for n in shape]), nrhs=eb.shape[0], depth=depth)
    out = s.solve(lam, mu, eb, rtol=rtol, maxiter=maxiter, want_sigma=True)
    s.close()
    return out


@pytest.mark.parametrize("first,second", [(((40, 36, 32), 4, 6), ((24, 20, 16), 0, 3))])
def test_noalloc_matches_the_allocating_solve_on_two_calls(gpu, first, second):
    for shape, depth, R in (first, second):
        lam, mu, _, _ = crossply_lame(shape, tuple(float(n) for n in shape), 2, 3.0, 0.4, seed=sum(shape))
        eb = np.random.default_rng(depth).standard_normal((R, 6)).astype(np.float32)
        print(f"noalloc {shape} depth {depth} load cases {R}", flush=True)
        rc, eps, sig, iters, res, smean = _noalloc(shape, lam, mu, eb, depth)
        ref = _reference(shape, lam, mu, eb, depth)
        assert rc == 0 and ref["converged"]
        assert np.array_equal(iters, ref["iters"]), (iters, ref["iters"])
        assert np.array_equal(smean, ref["sigma_mean"])
        assert torch.equal(eps, ref["epsilon"])
        assert torch.equal(sig, ref["sigma"])


def test_noalloc_repeats_and_refuses_oversized_problems(gpu, capfd):
    shape = (32, 32, 32)
    lam, mu, _, _ = crossply_lame(shape, (32.0,) * 3, 2, 3.0, 0.4, seed=5)
    eb = np.eye(6, dtype=np.float32)
    a = _noalloc(shape, lam, mu, eb, 2)
    b = _noalloc(shape, lam, mu, eb, 2)
    assert a[0] == 0 and b[0] == 0
    assert np.array_equal(a[3], b[3]) and np.array_equal(a[5], b[5]) and torch.equal(a[1], b[1])
    big = (256, 256, 256)
    lb = torch.ones(big, dtype=torch.float32, device=gpu)
    capfd.readouterr()
    rc = _noalloc(big, lb, lb, np.eye(6, dtype=np.float32)[:1], 0, want_sigma=False)[0]
    assert rc == -1
    assert "bytes of scratch" in capfd.readouterr().err
    # the array is still usable after the refusal
    c = _noalloc(shape, lam, mu, eb, 2)
    assert c[0] == 0 and np.array_equal(c[5], a[5])
