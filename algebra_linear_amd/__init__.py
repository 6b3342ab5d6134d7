"""algebra_linear_amd: the dense BLAS-3 / LAPACK solve path of the algebra_linear crate on
one MI355X (gfx950), as hand-written HIP behind a C ABI (include/algebra_linear_hip.h).

``blas`` and ``lapack`` mirror the crate's ``algebra_linear::blas`` / ``::lapack`` functions
of that path.  Importing the package loads libalgebra_linear_hip.so and fails loudly if it has
not been built; there is no CPU fallback.
"""
from . import _lib

_lib.load()

from . import blas, lapack  # noqa: E402
from .errors import BAD_SHAPE, CONVERGENCE, SINGULAR, BlasPanic, DeviceError, LapackError  # noqa: E402


def device_count() -> int:
    return int(_lib.load().al_device_count())


def version() -> str:
    return _lib.load().al_version().decode()


__all__ = ["blas", "lapack", "LapackError", "BlasPanic", "DeviceError", "SINGULAR", "BAD_SHAPE", "CONVERGENCE",
           "device_count", "version"]
