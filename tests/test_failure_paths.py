"""Failure semantics that need no GPU: a missing library raises (as the
reference loader does, src/jaxmaterials/solver/lippmann_schwinger.py:285-290),
and the loader rejects wrong shapes and dtypes before any call into the
library (its ndpointer argument types, :292-305)."""

import numpy as np
import pytest

from jaxmaterials_amd.common import GridSpec


def test_missing_library_raises(monkeypatch):
    from jaxmaterials_amd import _lib

    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/liblippmannschwinger.so")
    monkeypatch.setattr(_lib, "LIB_NAME", "liblippmannschwinger_not_on_the_loader_path.so")
    with pytest.raises(RuntimeError, match="Unable to load HIP library"):
        _lib.load()


def test_product_entry_points_raise_without_the_library(monkeypatch):
    from jaxmaterials_amd import _lib
    from jaxmaterials_amd.solver.lippmann_schwinger import effective_tensor, lippmann_schwinger_cuda

    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/liblippmannschwinger.so")
    monkeypatch.setattr(_lib, "LIB_NAME", "liblippmannschwinger_not_on_the_loader_path.so")
    gs = GridSpec(4, 4, 4, 1.0, 1.0, 1.0)
    a = np.ones((4, 4, 4), np.float32)
    with pytest.raises(RuntimeError, match="Unable to load"):
        lippmann_schwinger_cuda(a, a, np.zeros(6, np.float32), gs)
    with pytest.raises(RuntimeError, match="Unable to load"):
        effective_tensor(a, a, gs)


def test_loader_rejects_wrong_shapes_and_dtypes():
    import ctypes

    from jaxmaterials_amd.solver.lippmann_schwinger import lippmann_schwinger_cuda

    gs = GridSpec(4, 4, 4, 1.0, 1.0, 1.0)
    a = np.ones((4, 4, 4), np.float32)
    with pytest.raises(ValueError):
        lippmann_schwinger_cuda(a[:3], a, np.zeros(6, np.float32), gs)
    with pytest.raises(ValueError):
        lippmann_schwinger_cuda(a, a, np.zeros(5, np.float32), gs)
    # float64 fields fail the ndpointer dtype check, as in the reference
    with pytest.raises(ctypes.ArgumentError):
        lippmann_schwinger_cuda(a.astype(np.float64), a, np.zeros(6, np.float32), gs)
