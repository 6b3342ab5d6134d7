"""CPU-side checks of the C ABI: the library loads, exports every entry point the header
declares, and rejects bad shapes with the crate's messages before touching a device.
No compute call is made here (no GPU in this environment)."""
import ctypes
import os
import re

import numpy as np
import pytest

from algebra_linear_amd import _lib

HEADER = open(_lib.HEADER).read()
DECLARED = sorted(set(re.findall(r"\b(al_\w+)\s*\(", re.sub(r"/\*.*?\*/", "", HEADER, flags=re.S))))


def test_library_exports_every_declared_entry_point():
    lib = _lib.load()
    assert len(DECLARED) >= 20
    for name in DECLARED:
        assert hasattr(lib, name), name


def test_python_signature_table_covers_the_header():
    assert set(DECLARED) <= set(_lib.SIGNATURES) | {"al_prof_enable", "al_prof_reset", "al_prof_query"}


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def test_blas_shape_errors_carry_the_crates_panic_text():
    lib = _lib.load()
    a, b, c = np.zeros(5), np.zeros(6), np.zeros(4)
    assert lib.al_dgemm(1.0, _dp(a), 5, _dp(b), 6, 0.0, _dp(c), 4, 2, 2, 3) == _lib.AL_BAD_SHAPE
    assert _lib.last_error() == "dgemm: a must be m*k"
    assert lib.al_dgemm_tn(1.0, _dp(b), 6, _dp(a), 5, 0.0, _dp(c), 4, 2, 2, 3) == _lib.AL_BAD_SHAPE
    assert _lib.last_error() == "dgemm_tn: b must be k*n"
    assert lib.al_dgemm_tt(1.0, _dp(b), 6, _dp(b), 6, 0.0, _dp(a), 5, 2, 2, 3) == _lib.AL_BAD_SHAPE
    assert _lib.last_error() == "dgemm_tt: c must be m*n"
    assert lib.al_dtrsm(1.0, _dp(c), 4, _dp(a), 5, 2, 2, 1, 1, 0) == _lib.AL_BAD_SHAPE
    assert _lib.last_error() == "dtrsm: b must be m*n"
    assert lib.al_dtrsm(1.0, _dp(a), 5, _dp(c), 4, 2, 2, 0, 1, 0) == _lib.AL_BAD_SHAPE
    assert _lib.last_error() == "dtrsm: a must be n*n for the right side"
    assert lib.al_dsyrk(1.0, _dp(a), 5, 0.0, _dp(c), 4, 2, 2, 0) == _lib.AL_BAD_SHAPE
    assert _lib.last_error() == "dsyrk: a must be n*k"


def test_lapack_shape_errors_are_badshape_and_empty_inputs_are_ok():
    lib = _lib.load()
    sp = ctypes.POINTER(ctypes.c_size_t)
    piv = np.zeros(4, dtype=np.uintp)
    a = np.zeros(3)
    det = ctypes.c_double(0.0)
    assert lib.al_dgetrf(_dp(a), 3, 2, piv.ctypes.data_as(sp), 2, ctypes.byref(det)) == _lib.AL_BAD_SHAPE
    assert lib.al_dgetrf_rect(_dp(a), 4, 1, 2, piv.ctypes.data_as(sp), 2) == _lib.AL_BAD_SHAPE  # nrows < ncols
    assert lib.al_dgesv(_dp(a), 3, _dp(a), 2) == _lib.AL_BAD_SHAPE
    assert lib.al_dpotrf(_dp(a), 3, 2) == _lib.AL_BAD_SHAPE
    # n = 0 never needs a device (src/lapack/mod.rs:609-616)
    assert lib.al_dgetrf(_dp(a), 0, 0, piv.ctypes.data_as(sp), 0, ctypes.byref(det)) == _lib.AL_OK
    assert det.value == 1.0
    assert lib.al_dgesv(_dp(a), 0, _dp(a), 0) == _lib.AL_OK
    assert lib.al_dpotrf(_dp(a), 0, 0) == _lib.AL_OK


def test_size_overflow_is_a_shape_error_not_a_wrap():
    lib = _lib.load()
    a = np.zeros(4)
    big = 1 << 40
    assert lib.al_dgemm(1.0, _dp(a), 4, _dp(a), 4, 0.0, _dp(a), 4, big, big, big) == _lib.AL_BAD_SHAPE


@pytest.mark.skipif(_lib.load().al_device_count() > 0, reason="a GPU is visible")
def test_without_a_device_compute_fails_loudly():
    import algebra_linear_amd as al
    with pytest.raises(al.DeviceError):
        al.blas.dgemm(1.0, np.ones(4), np.ones(4), 0.0, np.zeros(4), 2, 2, 2)


def test_python_mirror_raises_the_crates_error_kinds():
    import algebra_linear_amd as al
    with pytest.raises(al.BlasPanic, match="dgemm_nt: b must be n\\*k"):
        al.blas.dgemm_nt(1.0, np.zeros(6), np.zeros(5), 0.0, np.zeros(4), 2, 2, 3)
    with pytest.raises(al.LapackError) as e:
        al.lapack.dpotrf(np.zeros(3), 2)
    assert e.value.kind == al.BAD_SHAPE
    with pytest.raises(TypeError):
        al.lapack.dgesv([1.0, 2.0, 3.0, 4.0], np.zeros(2))  # outputs must be ndarrays (&mut [f64])


def test_complex_shape_errors_and_empty_inputs():
    # src/blas.rs:1024-1074 (panics with these texts), src/lapack/complex.rs:67-347 (BadShape)
    import algebra_linear_amd as al
    lib = _lib.load()
    C = _lib.Complex.of
    z = np.zeros(6, dtype=np.complex128)
    zp = z.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.al_zgemm(C(1), zp, 5, zp, 6, C(0), zp, 4, 2, 2, 3) == _lib.AL_BAD_SHAPE
    assert _lib.last_error() == "zgemm: a must be m*k"
    assert lib.al_zgemm_tn(C(1), zp, 6, zp, 5, C(0), zp, 4, 2, 2, 3) == _lib.AL_BAD_SHAPE
    assert _lib.last_error() == "zgemm_tn: b must be k*n"
    assert lib.al_zgemm_nt(C(1), zp, 6, zp, 5, C(0), zp, 4, 2, 2, 3) == _lib.AL_BAD_SHAPE
    assert _lib.last_error() == "zgemm_nt: b must be n*k"
    assert lib.al_zgemm_tt(C(1), zp, 6, zp, 6, C(0), zp, 5, 2, 2, 3) == _lib.AL_BAD_SHAPE
    assert _lib.last_error() == "zgemm_tt: c must be m*n"
    with pytest.raises(al.BlasPanic, match="zgemm_tn: a must be k\\*m"):
        al.blas.zgemm_tn(1.0, np.zeros(5, complex), np.zeros(6, complex), 0.0, np.zeros(4, complex), 2, 2, 3)
    for call in (lambda: al.lapack.zgetrf(np.zeros(3, complex), 2),
                 lambda: al.lapack.zgesv(np.zeros(3, complex), np.zeros(2, complex)),
                 lambda: al.lapack.zpotrf(np.zeros(3, complex), 2, False),
                 lambda: al.lapack.zgetrf_rect(np.zeros(4, complex), 1, 2, np.zeros(2, dtype=np.uintp))):
        with pytest.raises(al.LapackError) as e:
            call()
        assert e.value.kind == al.BAD_SHAPE
    # n = 0 is a no-op everywhere (src/lapack/complex.rs:611-619)
    assert al.lapack.zgetrf(np.zeros(0, complex), 0) == []
    al.lapack.zgesv(np.zeros(0, complex), np.zeros(0, complex))
    al.lapack.zpotrf(np.zeros(0, complex), 0, True)
    with pytest.raises(TypeError):
        al.lapack.zgesv(np.zeros(4, complex), np.zeros(2))  # b must be complex128 (&mut [Complex64])
