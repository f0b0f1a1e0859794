"""Host-memory entry points of the HIP Lippmann-Schwinger solver.

``lippmann_schwinger_cuda`` keeps the name, arguments, defaults and error
behaviour of the reference loader (src/jaxmaterials/solver/lippmann_schwinger.py:268-329)
so existing callers switch by changing the import.  As in the reference, it
passes ``mu`` in the C ``lambda`` slot and ``lmbda`` in the ``mu`` slot
(lippmann_schwinger.py:310-312 against cuda/src/lippmann_schwinger.hh:178);
datasets generated through it therefore hold the same (swapped-moduli)
results as the reference's.  The C function itself implements the header's
argument meaning: the first pointer is lambda.
"""

import ctypes

import numpy as np

from .. import _lib

__all__ = ["lippmann_schwinger_cuda", "effective_tensor"]


def lippmann_schwinger_cuda(
    lmbda, mu, epsilon_bar, grid_spec, rtol=1e-6, atol=1.0e-20, maxiter=32, verbose=0
):
    """Solve for the strain with mean epsilon_bar on the GPU through the C ABI
    of liblippmannschwinger.so (the in-tree build, else one on the loader path).

    lmbda, mu are float32 C-ordered fields of grid_spec's shape; epsilon_bar is
    the Voigt mean strain (6 entries, tensor shear).  The loop stops once the
    equilibrium residual, relative to the residual of the uniform start, falls
    below rtol or below atol; it runs at most maxiter strain updates.  verbose 0
    prints nothing, 1 a one-line summary, 2 the residual of every iteration.
    Returns (epsilon, sigma, iterations); raises RuntimeError when the library
    cannot be loaded, reports an error, or the iteration count reaches maxiter.
    """
    lib = _lib.load()
    nd = np.ctypeslib.ndpointer
    # argument types of the reference loader (lippmann_schwinger.py:292-305):
    # ndpointer checks dtype float32 / int32 and C contiguity
    proto = ctypes.CFUNCTYPE(
        ctypes.c_int,
        nd(ctypes.c_float, flags="C_CONTIGUOUS"),
        nd(ctypes.c_float, flags="C_CONTIGUOUS"),
        nd(ctypes.c_float, flags="C_CONTIGUOUS"),
        nd(ctypes.c_float, flags="C_CONTIGUOUS"),
        nd(ctypes.c_float, flags="C_CONTIGUOUS"),
        nd(ctypes.c_int, flags="C_CONTIGUOUS"),
        nd(ctypes.c_float, flags="C_CONTIGUOUS"),
        ctypes.c_float,
        ctypes.c_float,
        ctypes.c_int,
        ctypes.c_int,
    )
    code = proto(("lippmann_schwinger_solve", lib))
    cells = np.array([grid_spec.nx, grid_spec.ny, grid_spec.nz], dtype=np.int32)
    extents = np.array([grid_spec.Lx, grid_spec.Ly, grid_spec.Lz], dtype=np.float32)
    epsilon = np.empty((6, grid_spec.nx, grid_spec.ny, grid_spec.nz), dtype=np.float32)
    sigma = np.empty((6, grid_spec.nx, grid_spec.ny, grid_spec.nz), dtype=np.float32)
    n = grid_spec.nx * grid_spec.ny * grid_spec.nz
    if np.asarray(mu).size != n or np.asarray(lmbda).size != n or np.asarray(epsilon_bar).size != 6:
        # the reference does not check shapes and reads out of bounds instead
        raise ValueError("lmbda and mu need nx*ny*nz entries and epsilon_bar 6")
    it = code(
        np.asarray(mu),
        np.asarray(lmbda),
        np.asarray(epsilon_bar, dtype=np.float32),
        np.asarray(epsilon),
        np.asarray(sigma),
        cells,
        extents,
        rtol,
        atol,
        maxiter,
        verbose,
    )
    if it < 0:
        raise RuntimeError("HIP Lippmann Schwinger solver reported an error (see stderr)")
    if it == maxiter:
        raise RuntimeError(f"Solver failed to converge after {maxiter} iterations")
    return (
        epsilon,
        sigma,
        it,
    )


def effective_tensor(lmbda, mu, grid_spec, rtol=1e-6, atol=0.0, maxiter=1000, depth=0, verbose=0):
    """Effective 6x6 tensor: row k = <sigma> under epsilon_bar = e_k.

    One batched solve of the six load cases of the reference's C_eff loop
    (src/jaxmaterials/data/data.py:178-186), on the argument order of the C
    header (first lambda, then mu; no swap).  Returns (C_eff float64 [6,6],
    iterations [6]).  Raises RuntimeError if a load case does not converge.
    """
    lib = _lib.load()
    lm = np.ascontiguousarray(lmbda, dtype=np.float32)
    m = np.ascontiguousarray(mu, dtype=np.float32)
    cells = np.array([grid_spec.nx, grid_spec.ny, grid_spec.nz], dtype=np.int32)
    extents = np.array([grid_spec.Lx, grid_spec.Ly, grid_spec.Lz], dtype=np.float32)
    ceff = np.zeros((6, 6), dtype=np.float64)
    iters = np.zeros(6, dtype=np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    rc = lib.lippmann_schwinger_effective_tensor(
        p(lm), p(m), p(cells), p(extents), rtol, atol, maxiter, depth, verbose, p(ceff), p(iters)
    )
    if rc < 0:
        raise RuntimeError("HIP Lippmann Schwinger solver reported an error (see stderr)")
    if rc == 1:
        raise RuntimeError(f"Solver failed to converge after {maxiter} iterations")
    return ceff, iters
