"""Device-memory solver for PyTorch tensors (ROCm), over the C ABI of
liblippmannschwinger.so (``ls_solver_*`` in include/lippmann_schwinger.h).

Torch is used for device memory only; all arithmetic of the solve runs in the
library's HIP kernels.  There is no CPU fallback: a missing library or a tensor
that is not on the GPU raises.
"""

import ctypes

import numpy as np
import torch

from .. import _lib

__all__ = ["LSSolver", "unit_load_cases"]


def unit_load_cases():
    """The six unit Voigt strains e_0..e_5 (rows), float32 [6, 6]."""
    return np.eye(6, dtype=np.float32)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _check_dev(name, t, shape=None):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise RuntimeError(f"{name} must be a GPU tensor (no CPU path exists)")
    if t.dtype != torch.float32 or not t.is_contiguous():
        raise TypeError(f"{name} must be contiguous float32")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")


class LSSolver:
    """Batched solver: ``nrhs`` load cases over ``nmat`` Lame-field pairs on one grid.

    :arg grid_spec: GridSpec (nx, ny, nz, Lx, Ly, Lz)
    :arg nrhs: number of load cases solved together
    :arg nmat: number of (lambda, mu) pairs
    :arg depth: Anderson depth (0 = plain Lippmann-Schwinger iteration)
    """

    def __init__(self, grid_spec, nrhs=6, nmat=1, depth=0):
        self.lib = _lib.load()
        self.gs = grid_spec
        self.shape = (grid_spec.nx, grid_spec.ny, grid_spec.nz)
        self.nrhs, self.nmat, self.depth = int(nrhs), int(nmat), int(depth)
        self._vox = np.array(self.shape, dtype=np.int32)
        self._ext = np.array([grid_spec.Lx, grid_spec.Ly, grid_spec.Lz], dtype=np.float32)
        h = self.lib.ls_solver_create(
            self._vox.ctypes.data_as(ctypes.c_void_p), self._ext.ctypes.data_as(ctypes.c_void_p),
            self.nrhs, self.nmat, self.depth, None,
        )
        if not h:
            raise RuntimeError("ls_solver_create failed (see stderr)")
        self._h = ctypes.c_void_p(h)

    def close(self):
        if getattr(self, "_h", None):
            self.lib.ls_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, lmbda, mu, epsilon_bar, rhs_mat=None, tau=None, epsilon=None, warm_start=False,
              want_sigma=False, rtol=1e-6, atol=0.0, maxiter=1000, rel_ref=0.0, absolute_residual=False,
              verbose=0, check=True, graph=False, cg=False, compatible_start=False):
        """Solve all load cases.

        :arg lmbda, mu: GPU float32 [nmat, nx, ny, nz] (or [nx, ny, nz] when nmat == 1)
        :arg epsilon_bar: array-like [nrhs, 6] of macroscopic strains
        :arg rhs_mat: material index of each load case (default all 0)
        :arg tau: optional GPU float32 [nrhs, 6, nx, ny, nz] polarisation added to the stress
        :arg epsilon: optional GPU float32 [nrhs, 6, nx, ny, nz]; start field if warm_start
        :arg rel_ref: normaliser of rtol; 0 means the residual of the uniform start
        :arg absolute_residual: residual without the mean-stress normalisation
        :arg graph: replay the iterations between host synchronisations as a captured HIP graph
        :arg cg: conjugate gradients instead of the Lippmann-Schwinger / Anderson update (same
            equation, residual, stop test and iteration semantics; Hooke's law only)
        :arg compatible_start: the warm start is compatible with mean epsilon_bar already
            (its projection is skipped)
        Returns a dict with epsilon, sigma (or None), iters, residual, sigma_mean (float64 [nrhs, 6]).
        """
        R = self.nrhs
        field = (R, 6) + self.shape
        lm = lmbda.reshape((self.nmat,) + self.shape) if lmbda.dim() == 3 else lmbda
        m = mu.reshape((self.nmat,) + self.shape) if mu.dim() == 3 else mu
        _check_dev("lmbda", lm, (self.nmat,) + self.shape)
        _check_dev("mu", m, (self.nmat,) + self.shape)
        eb = np.ascontiguousarray(np.asarray(epsilon_bar, dtype=np.float32).reshape(R, 6))
        mats = np.zeros(R, dtype=np.int32) if rhs_mat is None else np.ascontiguousarray(rhs_mat, dtype=np.int32)
        if tau is not None:
            _check_dev("tau", tau, field)
        if epsilon is None:
            if warm_start:
                raise ValueError("warm_start needs epsilon")
            epsilon = torch.empty(field, dtype=torch.float32, device=lm.device)
        _check_dev("epsilon", epsilon, field)
        sigma = torch.empty(field, dtype=torch.float32, device=lm.device) if want_sigma else None
        iters = np.zeros(R, dtype=np.int32)
        res = np.zeros(R, dtype=np.float64)
        smean = np.zeros((R, 6), dtype=np.float64)
        torch.cuda.current_stream(lm.device).synchronize()
        rc = self.lib.ls_solver_solve(
            self._h, _ptr(lm), _ptr(m), mats.ctypes.data_as(ctypes.c_void_p), eb.ctypes.data_as(ctypes.c_void_p),
            _ptr(tau), _ptr(epsilon), _ptr(sigma), int(bool(warm_start)), float(rtol), float(atol), int(maxiter),
            float(rel_ref), (1 if absolute_residual else 0) | (2 if graph else 0) | (4 if cg else 0)
            | (8 if compatible_start else 0), int(verbose),
            iters.ctypes.data_as(ctypes.c_void_p),
            res.ctypes.data_as(ctypes.c_void_p), smean.ctypes.data_as(ctypes.c_void_p),
        )
        if rc < 0:
            raise RuntimeError("ls_solver_solve failed (see stderr)")
        if check and rc == 1:
            raise RuntimeError(f"Solver failed to converge after {maxiter} iterations")
        return {"epsilon": epsilon, "sigma": sigma, "iters": iters, "residual": res, "sigma_mean": smean,
                "converged": rc == 0}

    def set_damage(self, damage=None, k=1e-12):
        """Material law of the following solves: the phase-field damaged law
        (``ls_solver_set_damage``) with GPU float32 damage fields [nmat, nx, ny, nz]
        (or [nx, ny, nz] when nmat == 1) and stability constant k, or Hooke's
        law when damage is None.  The tensor is kept referenced while set."""
        if damage is None:
            self._damage = None
            rc = self.lib.ls_solver_set_damage(self._h, None, 0.0)
        else:
            dm = damage.reshape((self.nmat,) + self.shape) if damage.dim() == 3 else damage
            _check_dev("damage", dm, (self.nmat,) + self.shape)
            self._damage = dm
            rc = self.lib.ls_solver_set_damage(self._h, _ptr(dm), float(k))
        if rc != 0:
            raise RuntimeError("ls_solver_set_damage failed")

    def phase_field(self, history, d_old, gc, lc, tolerance=1e-6, maxiter=1000, out=None):
        """phase_field_solve on this solver's grid (``ls_solver_phase_field``):
        GPU float64 [nx, ny, nz] history and previous damage; returns
        (damage float64 tensor, iterations, last residual)."""
        for name, t in (("history", history), ("d_old", d_old)):
            if not t.is_cuda or t.dtype != torch.float64 or tuple(t.shape) != self.shape or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous float64 GPU tensor {self.shape}")
        if out is None:
            out = torch.empty_like(d_old)
        it = ctypes.c_int(0)
        res = ctypes.c_double(0.0)
        # the domain size in FP64, as the reference's FP64 phase field builds xi~
        ext = np.array([self.gs.Lx, self.gs.Ly, self.gs.Lz], dtype=np.float64)
        torch.cuda.current_stream(history.device).synchronize()
        rc = self.lib.ls_solver_phase_field(self._h, _ptr(history), _ptr(d_old), float(gc), float(lc), float(tolerance),
                                            int(maxiter), _ptr(out), ctypes.byref(it), ctypes.byref(res),
                                            ext.ctypes.data_as(ctypes.c_void_p))
        if rc != 0:
            raise RuntimeError("ls_solver_phase_field failed (see stderr)")
        return out, it.value, res.value

    def autotune(self, lmbda, mu, epsilon=None, reps=3):
        """Time the candidate FFT tiles of the three passes on this solver's
        buffers and keep the fastest (``ls_solver_autotune``).  epsilon, if
        given, is used (and overwritten) as the strain scratch.  Returns the
        tuner's report."""
        lm = lmbda.reshape((self.nmat,) + self.shape) if lmbda.dim() == 3 else lmbda
        m = mu.reshape((self.nmat,) + self.shape) if mu.dim() == 3 else mu
        _check_dev("lmbda", lm, (self.nmat,) + self.shape)
        _check_dev("mu", m, (self.nmat,) + self.shape)
        if epsilon is None:
            epsilon = torch.empty((self.nrhs, 6) + self.shape, dtype=torch.float32, device=lm.device)
        _check_dev("epsilon", epsilon, (self.nrhs, 6) + self.shape)
        buf = ctypes.create_string_buffer(1024)
        torch.cuda.current_stream(lm.device).synchronize()
        rc = self.lib.ls_solver_autotune(self._h, _ptr(lm), _ptr(m), _ptr(epsilon), int(reps), buf, len(buf))
        if rc != 0:
            raise RuntimeError("ls_solver_autotune failed (see stderr)")
        return buf.value.decode()

    KINDS = ("z", "y", "x", "other", "z_aa", "aa_gram", "z_cg")

    def last_timing(self):
        """Kernel milliseconds and launch counts of the last solve by pass kind
        (``LSSolver.KINDS``), recorded when the environment variable LS_TIMING
        was set; also the number of base re-evaluations."""
        nk = len(self.KINDS)
        ms = np.zeros(nk, dtype=np.float64)
        cnt = np.zeros(nk, dtype=np.int32)
        nreb = self.lib.ls_solver_last_timing(self._h, ms.ctypes.data_as(ctypes.c_void_p),
                                              cnt.ctypes.data_as(ctypes.c_void_p), nk)
        return dict(zip(self.KINDS, ms.tolist())), dict(zip(self.KINDS, cnt.tolist())), int(nreb)
