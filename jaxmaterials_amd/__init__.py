"""MI355X-native Moulinec-Suquet Lippmann-Schwinger FFT elasticity solver.

The compute path is liblippmannschwinger.so (hand-written HIP kernels for
gfx950, C++ host code, a C ABI).  This package holds its Python side: the
drop-in loader ``solver.lippmann_schwinger.lippmann_schwinger_cuda`` (same
contract as the reference loader), a device-memory solver for PyTorch, the
adjoint, the microstructure generator and the learned warm start.
"""

from .common import GridSpec  # noqa: F401

__all__ = ["GridSpec"]
