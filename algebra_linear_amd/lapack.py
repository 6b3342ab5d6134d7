"""``algebra_linear::lapack`` on the GPU: LU, solve and Cholesky (src/lapack/mod.rs).

Same names and argument meaning as the crate; where the crate returns
``Err(LapackError::..)`` these raise ``LapackError`` with that ``kind``.  Matrices are
1-D float64 column-major numpy arrays overwritten in place; pivots are 0-based.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .blas import _in, _out, _p, _zin, _zout, _zp
from .errors import BAD_SHAPE, BlasPanic, LapackError, check_lapack

_sp = ctypes.POINTER(ctypes.c_size_t)


def _piv_out(ipiv) -> np.ndarray:
    if not (isinstance(ipiv, np.ndarray) and ipiv.dtype == np.uintp and ipiv.flags.c_contiguous):
        raise TypeError("ipiv must be a contiguous numpy array of dtype uintp (usize)")
    return ipiv.reshape(-1)


def dgetrf(a: np.ndarray, n: int):
    """In-place P*A = L*U; returns ``(ipiv, det)`` (src/lapack/mod.rs:181)."""
    aa = _out(a)
    ipiv = np.zeros(n, dtype=np.uintp)
    det = ctypes.c_double(0.0)
    check_lapack(_lib.load().al_dgetrf(_p(aa), aa.size, n, ipiv.ctypes.data_as(_sp), ipiv.size, ctypes.byref(det)))
    return [int(v) for v in ipiv], det.value


def dgetrf_rect(a: np.ndarray, nrows: int, ncols: int, ipiv: np.ndarray) -> None:
    """Panel LU of nrows x ncols, nrows >= ncols; ``ipiv`` (len ncols) filled (src/lapack/mod.rs:202)."""
    aa, pv = _out(a), _piv_out(ipiv)
    check_lapack(_lib.load().al_dgetrf_rect(_p(aa), aa.size, nrows, ncols, pv.ctypes.data_as(_sp), pv.size))


def dgetrs(lu, ipiv, b: np.ndarray) -> None:
    """Solve with a ``dgetrf`` factorization; ``n = len(b)`` (src/lapack/mod.rs:330)."""
    lu = _in(lu)
    bb = _out(b)
    pv = np.ascontiguousarray(np.asarray(ipiv, dtype=np.uintp))
    check_lapack(_lib.load().al_dgetrs(_p(lu), lu.size, pv.ctypes.data_as(_sp), pv.size, _p(bb), bb.size))


def dgesv(a: np.ndarray, b: np.ndarray) -> None:
    """Solve A x = b for one right-hand side, ``n = len(b)``; ``a`` is destroyed (src/lapack/mod.rs:309)."""
    aa, bb = _out(a), _out(b)
    check_lapack(_lib.load().al_dgesv(_p(aa), aa.size, _p(bb), bb.size))


def dgesv_nrhs(a: np.ndarray, b: np.ndarray, n: int, nrhs: int) -> list:
    """Extension: A X = B with B n x nrhs column-major; returns the pivots."""
    aa, bb = _out(a), _out(b)
    ipiv = np.zeros(max(n, 1), dtype=np.uintp)
    check_lapack(_lib.load().al_dgesv_nrhs(_p(aa), aa.size, _p(bb), bb.size, n, nrhs, ipiv.ctypes.data_as(_sp)))
    return [int(v) for v in ipiv[:n]]


def dpotrf(a: np.ndarray, n: int) -> None:
    """Lower Cholesky in place; the upper triangle is neither read nor written (src/lapack/mod.rs:368)."""
    aa = _out(a)
    check_lapack(_lib.load().al_dpotrf(_p(aa), aa.size, n))


def dgetrs_nrhs(lu, ipiv, b: np.ndarray, n: int, nrhs: int) -> None:
    """Extension: solve with a ``dgetrf`` factorization for B n x nrhs column-major, in place."""
    lu = _in(lu)
    bb = _out(b)
    pv = np.ascontiguousarray(np.asarray(ipiv, dtype=np.uintp))
    check_lapack(_lib.load().al_dgetrs_nrhs(_p(lu), lu.size, pv.ctypes.data_as(_sp), pv.size, _p(bb), bb.size, n,
                                            nrhs))


def dgetri(lu: np.ndarray, n: int, ipiv) -> None:
    """In-place inverse from a ``dgetrf`` factorization (src/lapack/mod.rs:249); raises ``Singular``
    (``lu`` unchanged) when U has a zero on its diagonal."""
    aa = _out(lu)
    pv = np.ascontiguousarray(np.asarray(ipiv, dtype=np.uintp))
    check_lapack(_lib.load().al_dgetri(_p(aa), aa.size, n, pv.ctypes.data_as(_sp), pv.size))


def dsyr2k(upper: bool, m: int, k: int, alpha: float, v, w, beta: float, c: np.ndarray) -> None:
    """dsytrd's symmetric rank-2k update (src/lapack/sym.rs:582): one triangle of
    C := alpha*V*W^T + alpha*W*V^T + beta*C with V, W m x k and C m x m, packed column-major."""
    vv, ww, cc = _in(v), _in(w), _out(c)
    st = _lib.load().al_dsyr2k(int(upper), m, k, float(alpha), _p(vv), vv.size, _p(ww), ww.size, float(beta), _p(cc),
                               cc.size)
    if st == _lib.AL_BAD_SHAPE:
        raise BlasPanic(_lib.last_error())
    check_lapack(st)


__all__ = ["dgetrf", "dgetrf_rect", "dgetrs", "dgetrs_nrhs", "dgetri", "dgesv", "dgesv_nrhs", "dpotrf", "dsyr2k",
           "LapackError", "BAD_SHAPE"]


FACTORS = {"fp64": _lib.AL_FACTOR_FP64, "fp32": _lib.AL_FACTOR_FP32, "bf16": _lib.AL_FACTOR_BF16}


def ir_options(max_iter: int = 30, allow_fallback: bool = True, tol: float = 1e-14,
               stop_tol: float = 1e-15) -> "_lib.IrOptions":
    """Refinement controls of :func:`dgesv_mixed` (the C defaults when called without arguments)."""
    return _lib.IrOptions(int(max_iter), 1 if allow_fallback else 0, float(tol), float(stop_tol))


def dgesv_mixed(a, b: np.ndarray, n: int, nrhs: int = 1, factor: str = "fp32", options=None) -> dict:
    """Extension: solve A X = B (B n x nrhs, overwritten by X) with a low-precision pivoted
    LU and FP64 iterative refinement, falling back to the FP64 pivoted LU; ``a`` is not
    modified.  Returns the refinement statistics as a dict; raises ``LapackError`` with kind
    ``Convergence`` if refinement fails and the fallback is disabled."""
    aa, bb = _in(a), _out(b)
    if factor not in FACTORS:
        raise ValueError(f"factor must be one of {sorted(FACTORS)}")
    opt = options if options is not None else ir_options()
    st = _lib.IrStats()
    check_lapack(_lib.load().al_dgesv_mixed(_p(aa), aa.size, _p(bb), bb.size, n, nrhs, FACTORS[factor],
                                            ctypes.byref(opt), ctypes.byref(st)))
    return {name: getattr(st, name) for name, _ in _lib.IrStats._fields_}


# --- complex twins (src/lapack/complex.rs): complex128 arrays -----------------------------

def zgetrf(a: np.ndarray, n: int) -> list:
    """In-place P*A = L*U of a complex n x n matrix, |Re|+|Im| pivoting; returns the pivots, as the
    crate's ``Vec<usize>`` (src/lapack/complex.rs:67)."""
    aa = _zout(a)
    ipiv = np.zeros(max(n, 1), dtype=np.uintp)
    check_lapack(_lib.load().al_zgetrf(_zp(aa), aa.size, n, ipiv.ctypes.data_as(_sp), n))
    return [int(v) for v in ipiv[:n]]


def zgetrf_rect(a: np.ndarray, nrows: int, ncols: int, ipiv: np.ndarray) -> None:
    """Complex panel LU, nrows >= ncols; ``ipiv`` (len ncols) filled (src/lapack/complex.rs:169)."""
    aa, pv = _zout(a), _piv_out(ipiv)
    check_lapack(_lib.load().al_zgetrf_rect(_zp(aa), aa.size, nrows, ncols, pv.ctypes.data_as(_sp), pv.size))


def zgesv(a: np.ndarray, b: np.ndarray) -> None:
    """Solve A x = b (one right-hand side, n = len(b)); ``a`` destroyed, ``b`` := x
    (src/lapack/complex.rs:279)."""
    aa, bb = _zout(a), _zout(b)
    check_lapack(_lib.load().al_zgesv(_zp(aa), aa.size, _zp(bb), bb.size))


def zpotrf(a: np.ndarray, n: int, upper: bool) -> None:
    """Complex Cholesky, A = L L^H (lower) or U^H U (upper); only that triangle is read or written
    (src/lapack/complex.rs:344)."""
    aa = _zout(a)
    check_lapack(_lib.load().al_zpotrf(_zp(aa), aa.size, n, int(upper)))
