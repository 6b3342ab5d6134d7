"""CPU oracle for the dense solve path -- TEST INFRASTRUCTURE ONLY.

A ctypes view of ``oracle/liboracle.so``, the plain-C restatement of the
algebra_linear crate's own CPU routines (``oracle/al_oracle.c``; every C function
cites the reference file:line it follows).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg use it, and
only as the checker or the timed CPU baseline; the product library in
``algebra_linear_amd`` never imports it.

Arrays are numpy float64 in column-major order flattened to 1-D (element (i, j) of
an ``nrow``-row matrix at ``[i + j*nrow]``), exactly the crate's slice layout.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_size_p = ctypes.POINTER(ctypes.c_size_t)
_c_u32_p = ctypes.POINTER(ctypes.c_uint32)


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L: ctypes.CDLL) -> None:
    d, sz, i, ch = ctypes.c_double, ctypes.c_size_t, ctypes.c_int, ctypes.c_char
    sig = {
        "or_idamax": (sz, [_c_double_p, sz]),
        "or_dnrm2": (d, [_c_double_p, sz]),
        "or_dgemv": (i, [d, _c_double_p, _c_double_p, d, _c_double_p, sz, sz]),
        "or_dgemv_trans": (i, [d, _c_double_p, _c_double_p, d, _c_double_p, sz, sz]),
        "or_dlange": (d, [ch, _c_double_p, sz, sz]),
        "or_set_avx": (None, [i]),
        "or_gemm": (i, [i, i, d, _c_double_p, _c_double_p, d, _c_double_p, sz, sz, sz]),
        "or_dsyrk": (i, [d, _c_double_p, d, _c_double_p, sz, sz, i]),
        "or_dsyrk_rank": (i, [d, _c_double_p, d, _c_double_p, sz, sz, i, i]),
        "or_dsyr2k": (None, [i, sz, sz, d, _c_double_p, sz, _c_double_p, sz, d, _c_double_p, sz]),
        "or_dtrsv": (i, [_c_double_p, _c_double_p, sz, i, i]),
        "or_dtrsv_transpose": (i, [_c_double_p, _c_double_p, sz, i, i]),
        "or_dtrsm": (i, [d, _c_double_p, _c_double_p, sz, sz, i, i, i]),
        "or_dgetrf": (i, [_c_double_p, sz, _c_size_p, _c_double_p]),
        "or_dgetrf_rect": (i, [_c_double_p, sz, sz, _c_size_p]),
        "or_dgetri": (i, [_c_double_p, sz, _c_size_p]),
        "or_dgetrs": (None, [_c_double_p, _c_size_p, _c_double_p, sz]),
        "or_dgesv": (i, [_c_double_p, _c_double_p, sz]),
        "or_dpotrf": (i, [_c_double_p, sz]),
        "or_get01": (d, [_c_double_p, _c_double_p, _c_size_p, sz, sz]),
        "or_get03": (d, [_c_double_p, _c_double_p, sz, _c_double_p]),
        "or_get04": (d, [_c_double_p, _c_double_p, sz, sz, d]),
        "or_pot03": (d, [_c_double_p, _c_double_p, sz, _c_double_p]),
        "or_laruv": (None, [_c_u32_p, sz, _c_double_p]),
        "or_larnv": (None, [i, _c_u32_p, sz, _c_double_p]),
        "or_laran": (d, [_c_u32_p]),
        "or_latm1": (i, [i, d, i, i, _c_u32_p, _c_double_p, sz]),
        "or_dlatms": (i, [i, i, i, d, d, i, sz, sz, _c_double_p, _c_double_p, _c_u32_p]),
        "or_backward_error": (d, [_c_double_p, _c_double_p, _c_double_p, sz, sz]),
        "or_zgemm": (i, [i, i, _c_double_p, _c_double_p, _c_double_p, _c_double_p, _c_double_p, sz, sz, sz]),
        "or_zgetrf": (i, [_c_double_p, sz, _c_size_p]),
        "or_zgetrf_rect": (i, [_c_double_p, sz, sz, _c_size_p]),
        "or_zgesv": (i, [_c_double_p, _c_double_p, sz]),
        "or_zpotrf": (i, [_c_double_p, sz, i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_c_double_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def _seed(seed) -> np.ndarray:
    s = np.ascontiguousarray(np.asarray(seed, dtype=np.uint32))
    assert s.shape == (4,)
    return s


# --- distribution codes of DLARNV -----------------------------------------------------------
UNIFORM, SYMMETRIC, NORMAL = 1, 2, 3
# --- DLATMS symmetry / packing ----------------------------------------------------------------
NONSYMMETRIC, POSITIVE_SYMMETRIC, SYMMETRIC_SYM = 0, 1, 2
FULL, UPPER, LOWER = 0, 1, 2


def idamax(x) -> int:
    x = _f64(x)
    return int(lib().or_idamax(_dp(x), x.size))


def dnrm2(x) -> float:
    x = _f64(x)
    return float(lib().or_dnrm2(_dp(x), x.size))


def dlange(norm: str, a, m: int, n: int) -> float:
    a = _f64(a)
    return float(lib().or_dlange(norm.encode()[:1], _dp(a), m, n))


def dgemv(alpha, a, x, beta, y, m, n) -> np.ndarray:
    a, x, y = _f64(a), _f64(x), _f64(y).copy()
    lib().or_dgemv(alpha, _dp(a), _dp(x), beta, _dp(y), m, n)
    return y


def dgemv_trans(alpha, a, x, beta, y, m, n) -> np.ndarray:
    a, x, y = _f64(a), _f64(x), _f64(y).copy()
    lib().or_dgemv_trans(alpha, _dp(a), _dp(x), beta, _dp(y), m, n)
    return y


def gemm(ta: bool, tb: bool, alpha, a, b, beta, c, m, n, k) -> np.ndarray:
    """``gemm_core`` of src/blas.rs:283; returns the new C."""
    a, b, c = _f64(a), _f64(b), _f64(c).copy()
    assert a.size >= m * k and b.size >= k * n and c.size >= m * n
    st = lib().or_gemm(int(ta), int(tb), alpha, _dp(a), _dp(b), beta, _dp(c), m, n, k)
    assert st == 0
    return c


def dgemm(alpha, a, b, beta, c, m, n, k):
    return gemm(False, False, alpha, a, b, beta, c, m, n, k)


def dgemm_tn(alpha, a, b, beta, c, m, n, k):
    return gemm(True, False, alpha, a, b, beta, c, m, n, k)


def dgemm_nt(alpha, a, b, beta, c, m, n, k):
    return gemm(False, True, alpha, a, b, beta, c, m, n, k)


def dgemm_tt(alpha, a, b, beta, c, m, n, k):
    return gemm(True, True, alpha, a, b, beta, c, m, n, k)


def set_avx(flag: int) -> None:
    lib().or_set_avx(int(flag))


def dsyrk(alpha, a, beta, c, n, k, upper) -> np.ndarray:
    a, c = _f64(a), _f64(c).copy()
    lib().or_dsyrk(alpha, _dp(a), beta, _dp(c), n, k, int(upper))
    return c


def dsyrk_rank(alpha, a, beta, c, n, k, upper, transposed) -> np.ndarray:
    a, c = _f64(a), _f64(c).copy()
    lib().or_dsyrk_rank(alpha, _dp(a), beta, _dp(c), n, k, int(upper), int(transposed))
    return c


def dsyr2k(upper, m, k, alpha, v, w, beta, c) -> np.ndarray:
    """dsytrd's dsyr2k (src/lapack/sym.rs:582-668) on packed V, W (m x k) and C (m x m)."""
    v, w, c = _f64(v), _f64(w), _f64(c).copy()
    lib().or_dsyr2k(int(upper), m, k, alpha, _dp(v), max(m, 1), _dp(w), max(m, 1), beta, _dp(c), max(m, 1))
    return c


def dtrsv(a, x, n, lower, unit) -> np.ndarray:
    a, x = _f64(a), _f64(x).copy()
    lib().or_dtrsv(_dp(a), _dp(x), n, int(lower), int(unit))
    return x


def dtrsv_transpose(a, x, n, lower, unit) -> np.ndarray:
    a, x = _f64(a), _f64(x).copy()
    lib().or_dtrsv_transpose(_dp(a), _dp(x), n, int(lower), int(unit))
    return x


def dtrsm(alpha, a, b, m, n, left, lower, unit) -> np.ndarray:
    a, b = _f64(a), _f64(b).copy()
    lib().or_dtrsm(alpha, _dp(a), _dp(b), m, n, int(left), int(lower), int(unit))
    return b


def dgetrf(a, n):
    """Returns (status, lu, ipiv, det); status 0 Ok, 1 Singular."""
    a = _f64(a).copy()
    ipiv = np.zeros(max(n, 1), dtype=np.uintp)
    det = ctypes.c_double(1.0)
    st = lib().or_dgetrf(_dp(a), n, ipiv.ctypes.data_as(_c_size_p), ctypes.byref(det))
    return st, a, ipiv[:n].astype(np.int64), det.value


def dgetrf_rect(a, nrows, ncols):
    a = _f64(a).copy()
    ipiv = np.zeros(max(ncols, 1), dtype=np.uintp)
    st = lib().or_dgetrf_rect(_dp(a), nrows, ncols, ipiv.ctypes.data_as(_c_size_p))
    return st, a, ipiv[:ncols].astype(np.int64)


def dgetri(lu, n, ipiv):
    lu = _f64(lu).copy()
    piv = np.ascontiguousarray(np.asarray(ipiv, dtype=np.uintp))
    st = lib().or_dgetri(_dp(lu), n, piv.ctypes.data_as(_c_size_p))
    return st, lu


def dgetrs(lu, ipiv, b) -> np.ndarray:
    lu, b = _f64(lu), _f64(b).copy()
    piv = np.ascontiguousarray(np.asarray(ipiv, dtype=np.uintp))
    lib().or_dgetrs(_dp(lu), piv.ctypes.data_as(_c_size_p), _dp(b), b.size)
    return b


def dgesv(a, b):
    a, b = _f64(a).copy(), _f64(b).copy()
    st = lib().or_dgesv(_dp(a), _dp(b), b.size)
    return st, a, b


def dpotrf(a, n):
    a = _f64(a).copy()
    st = lib().or_dpotrf(_dp(a), n)
    return st, a


def get01(a, lu, ipiv, m, n) -> float:
    a, lu = _f64(a), _f64(lu)
    piv = np.ascontiguousarray(np.asarray(ipiv, dtype=np.uintp))
    return float(lib().or_get01(_dp(a), _dp(lu), piv.ctypes.data_as(_c_size_p), m, n))


def get03(a, ainv, n):
    a, ainv = _f64(a), _f64(ainv)
    rc = ctypes.c_double(0.0)
    r = lib().or_get03(_dp(a), _dp(ainv), n, ctypes.byref(rc))
    return float(r), rc.value


def get04(x, xact, n, nrhs, rcond) -> float:
    x, xact = _f64(x), _f64(xact)
    return float(lib().or_get04(_dp(x), _dp(xact), n, nrhs, rcond))


def pot03(a, ainv, n):
    a, ainv = _f64(a), _f64(ainv)
    rc = ctypes.c_double(0.0)
    r = lib().or_pot03(_dp(a), _dp(ainv), n, ctypes.byref(rc))
    return float(r), rc.value


def laruv(seed, n):
    s = _seed(seed)
    x = np.zeros(max(n, 1))
    lib().or_laruv(s.ctypes.data_as(_c_u32_p), n, _dp(x))
    return x[:n], s


def larnv(dist: int, seed, n: int):
    s = _seed(seed)
    x = np.zeros(max(n, 1))
    lib().or_larnv(dist, s.ctypes.data_as(_c_u32_p), n, _dp(x))
    return x[:n], s


def latm1(mode, cond, irsign, dist, seed, d):
    s = _seed(seed)
    d = _f64(d).copy()
    info = lib().or_latm1(mode, cond, irsign, dist, s.ctypes.data_as(_c_u32_p), _dp(d), d.size)
    return info, d, s


def dlatms(dist, sym, mode, cond, dmax, pack, m, n, seed, d=None):
    s = _seed(seed)
    mn = min(m, n)
    d = np.zeros(max(mn, 1)) if d is None else _f64(d).copy()
    a = np.zeros(max(m * n, 1))
    info = lib().or_dlatms(dist, sym, mode, cond, dmax, pack, m, n, _dp(a), _dp(d), s.ctypes.data_as(_c_u32_p))
    return info, a[: m * n], d[:mn], s


def backward_error(a, x, b, n, nrhs=1) -> float:
    a, x, b = _f64(a), _f64(x), _f64(b)
    return float(lib().or_backward_error(_dp(a), _dp(x), _dp(b), n, nrhs))


# --- complex twins: numpy complex128 arrays (interleaved re, im, exactly Complex64's layout) ---

def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128).ravel())


def _zp(a: np.ndarray):
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.ctypes.data_as(_c_double_p)


def _zscalar(v) -> np.ndarray:
    v = complex(v)
    return np.array([v.real, v.imag], dtype=np.float64)


def zgemm(ta: bool, tb: bool, alpha, a, b, beta, c, m, n, k) -> np.ndarray:
    """``zgemm_core`` of src/blas.rs:1086-1118 (``zgemm``/``_tn``/``_nt``/``_tt``); returns the new C."""
    a, b, c = _c128(a), _c128(b), _c128(c).copy()
    assert a.size == m * k and b.size == k * n and c.size == m * n
    al, be = _zscalar(alpha), _zscalar(beta)
    lib().or_zgemm(int(ta), int(tb), _dp(al), _zp(a), _zp(b), _dp(be), _zp(c), m, n, k)
    return c


def zgetrf(a, n):
    """src/lapack/complex.rs:67-121: (status, LU, ipiv); status 0 Ok, 1 Singular."""
    a = _c128(a).copy()
    assert a.size == n * n
    ipiv = np.zeros(max(n, 1), dtype=np.uintp)
    st = lib().or_zgetrf(_zp(a), n, ipiv.ctypes.data_as(_c_size_p))
    return st, a, ipiv[:n].astype(np.int64)


def zgetrf_rect(a, nrows, ncols):
    """src/lapack/complex.rs:169-208: (status, LU, ipiv)."""
    a = _c128(a).copy()
    assert a.size == nrows * ncols
    ipiv = np.zeros(max(ncols, 1), dtype=np.uintp)
    st = lib().or_zgetrf_rect(_zp(a), nrows, ncols, ipiv.ctypes.data_as(_c_size_p))
    return st, a, ipiv[:ncols].astype(np.int64)


def zgesv(a, b):
    """src/lapack/complex.rs:279-320: (status, x); one right-hand side."""
    b = _c128(b).copy()
    n = b.size
    a = _c128(a).copy()
    assert a.size == n * n
    st = lib().or_zgesv(_zp(a), _zp(b), n)
    return st, b


def zpotrf(a, n, upper: bool):
    """src/lapack/complex.rs:344-400: (status, factor); only the requested triangle is written."""
    a = _c128(a).copy()
    assert a.size == n * n
    st = lib().or_zpotrf(_zp(a), n, int(upper))
    return st, a
