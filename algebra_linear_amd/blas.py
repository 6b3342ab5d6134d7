"""``algebra_linear::blas`` on the GPU: the level-3 routines of the dense solve path.

Same names, argument order and meaning as the crate's Rust functions (src/blas.rs):
slices are 1-D float64 numpy arrays in column-major order, outputs are written in place
(the crate's ``&mut [f64]``), and a shape mismatch raises ``BlasPanic`` with the crate's
panic message.  Every call runs on the MI355X through libalgebra_linear_hip.so.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .errors import BlasPanic, check_blas

_dp = ctypes.POINTER(ctypes.c_double)


def _in(x) -> np.ndarray:
    """A read-only operand: any float64-convertible sequence, flattened."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64)).reshape(-1)
    return a


def _out(x) -> np.ndarray:
    """An in-place operand: must already be a contiguous float64 ndarray (``&mut [f64]``)."""
    if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous and x.flags.writeable):
        raise TypeError("in-place operand must be a writeable, contiguous float64 numpy array")
    return x.reshape(-1)


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _gemm(fn, alpha, a, b, beta, c, m, n, k) -> None:
    a, b, cc = _in(a), _in(b), _out(c)
    check_blas(fn(float(alpha), _p(a), a.size, _p(b), b.size, float(beta), _p(cc), cc.size, m, n, k))


def dgemm(alpha: float, a, b, beta: float, c: np.ndarray, m: int, n: int, k: int) -> None:
    """C := alpha*A*B + beta*C; ``a`` m x k, ``b`` k x n, ``c`` m x n (src/blas.rs:201)."""
    _gemm(_lib.load().al_dgemm, alpha, a, b, beta, c, m, n, k)


def dgemm_tn(alpha: float, a, b, beta: float, c: np.ndarray, m: int, n: int, k: int) -> None:
    """C := alpha*A^T*B + beta*C; ``a`` k x m, ``b`` k x n (src/blas.rs:220)."""
    _gemm(_lib.load().al_dgemm_tn, alpha, a, b, beta, c, m, n, k)


def dgemm_nt(alpha: float, a, b, beta: float, c: np.ndarray, m: int, n: int, k: int) -> None:
    """C := alpha*A*B^T + beta*C; ``a`` m x k, ``b`` n x k (src/blas.rs:237)."""
    _gemm(_lib.load().al_dgemm_nt, alpha, a, b, beta, c, m, n, k)


def dgemm_tt(alpha: float, a, b, beta: float, c: np.ndarray, m: int, n: int, k: int) -> None:
    """C := alpha*A^T*B^T + beta*C; ``a`` k x m, ``b`` n x k (src/blas.rs:252)."""
    _gemm(_lib.load().al_dgemm_tt, alpha, a, b, beta, c, m, n, k)


def dtrsm(alpha: float, a, b: np.ndarray, m: int, n: int, left: bool, lower: bool, unit: bool) -> None:
    """B := alpha*A^-1*B (left, ``a`` m x m) or alpha*B*A^-1 (right, ``a`` n x n) (src/blas.rs:1384)."""
    a, bb = _in(a), _out(b)
    check_blas(_lib.load().al_dtrsm(float(alpha), _p(a), a.size, _p(bb), bb.size, m, n, int(left), int(lower),
                                    int(unit)))


def dsyrk(alpha: float, a, beta: float, c: np.ndarray, n: int, k: int, upper: bool) -> None:
    """One triangle of C := alpha*A*A^T + beta*C; ``a`` n x k (src/blas.rs:609)."""
    a, cc = _in(a), _out(c)
    check_blas(_lib.load().al_dsyrk(float(alpha), _p(a), a.size, float(beta), _p(cc), cc.size, n, k, int(upper)))


def dgemv(alpha: float, a, x, beta: float, y: np.ndarray, m: int, n: int) -> None:
    """y := alpha*A*x + beta*y; ``a`` m x n, ``x`` n, ``y`` m (src/blas.rs:102); bit for bit the crate's."""
    a, x, yy = _in(a), _in(x), _out(y)
    check_blas(_lib.load().al_dgemv(float(alpha), _p(a), a.size, _p(x), x.size, float(beta), _p(yy), yy.size, m, n))


def dgemv_trans(alpha: float, a, x, beta: float, y: np.ndarray, m: int, n: int) -> None:
    """y := alpha*A^T*x + beta*y; ``a`` m x n, ``x`` m, ``y`` n (src/blas.rs:144); bit for bit the crate's."""
    a, x, yy = _in(a), _in(x), _out(y)
    check_blas(_lib.load().al_dgemv_trans(float(alpha), _p(a), a.size, _p(x), x.size, float(beta), _p(yy), yy.size,
                                          m, n))


def dlange(norm: str, a, m: int, n: int) -> float:
    """|A| for ``norm`` in 1/O, I, F/E, M (src/blas.rs:638); an unknown norm panics, as in the crate."""
    a = _in(a)
    out = ctypes.c_double(0.0)
    st = _lib.load().al_dlange(str(norm).encode()[:1] or b"?", _p(a), a.size, m, n, ctypes.byref(out))
    if st == _lib.AL_ERR_ARG and _lib.last_error().startswith("dlange: unrecognized norm"):
        raise BlasPanic(_lib.last_error())
    check_blas(st)
    return out.value


def dtrsv(a, x: np.ndarray, n: int, lower: bool, unit: bool) -> None:
    """Solves A x = b in place, ``a`` n x n triangular (src/blas.rs:1355)."""
    a, xx = _in(a), _out(x)
    check_blas(_lib.load().al_dtrsv(_p(a), a.size, _p(xx), xx.size, n, int(lower), int(unit)))


def dtrsv_transpose(a, x: np.ndarray, n: int, lower: bool, unit: bool) -> None:
    """Solves A^T x = b in place (src/blas.rs:1366)."""
    a, xx = _in(a), _out(x)
    check_blas(_lib.load().al_dtrsv_transpose(_p(a), a.size, _p(xx), xx.size, n, int(lower), int(unit)))


# --- complex twins (src/blas.rs:1023-1118): complex128 arrays, Complex64's layout ----------

def _zin(x) -> np.ndarray:
    """A read-only complex operand, flattened (complex128 = interleaved re, im doubles)."""
    return np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).reshape(-1)


def _zout(x) -> np.ndarray:
    """An in-place complex operand (``&mut [Complex64]``)."""
    if not (isinstance(x, np.ndarray) and x.dtype == np.complex128 and x.flags.c_contiguous and x.flags.writeable):
        raise TypeError("in-place operand must be a writeable, contiguous complex128 numpy array")
    return x.reshape(-1)


def _zp(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _zgemm(fn, alpha, a, b, beta, c, m, n, k) -> None:
    a, b, cc = _zin(a), _zin(b), _zout(c)
    C = _lib.Complex.of
    check_blas(fn(C(alpha), _zp(a), a.size, _zp(b), b.size, C(beta), _zp(cc), cc.size, m, n, k))


def zgemm(alpha, a, b, beta, c: np.ndarray, m: int, n: int, k: int) -> None:
    """C := alpha*A*B + beta*C for complex data; ``a`` m x k, ``b`` k x n (src/blas.rs:1023)."""
    _zgemm(_lib.load().al_zgemm, alpha, a, b, beta, c, m, n, k)


def zgemm_tn(alpha, a, b, beta, c: np.ndarray, m: int, n: int, k: int) -> None:
    """C := alpha*A^T*B + beta*C, not conjugated; ``a`` k x m (src/blas.rs:1042)."""
    _zgemm(_lib.load().al_zgemm_tn, alpha, a, b, beta, c, m, n, k)


def zgemm_nt(alpha, a, b, beta, c: np.ndarray, m: int, n: int, k: int) -> None:
    """C := alpha*A*B^T + beta*C, not conjugated; ``b`` n x k (src/blas.rs:1059)."""
    _zgemm(_lib.load().al_zgemm_nt, alpha, a, b, beta, c, m, n, k)


def zgemm_tt(alpha, a, b, beta, c: np.ndarray, m: int, n: int, k: int) -> None:
    """C := alpha*A^T*B^T + beta*C; ``a`` k x m, ``b`` n x k (src/blas.rs:1071)."""
    _zgemm(_lib.load().al_zgemm_tt, alpha, a, b, beta, c, m, n, k)
