"""Device-pointer API on HBM buffers (torch used only for device memory and streams)."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
al = pytest.importorskip("algebra_linear_amd")
from algebra_linear_amd import _lib  # noqa: E402

LIB = _lib.load()


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def seq_swaps(a2d, ipiv, k0, k1):
    a = a2d.copy()
    for s in range(k0, k1):
        p = ipiv[s]
        a[[s, p], :] = a[[p, s], :]
    return a


@pytest.mark.parametrize("case", ["random", "repeat_outside", "chain", "identity", "long"])
def test_dlaswp_equals_the_sequential_swaps(case):
    dev = _dev()
    rng = np.random.default_rng(5)
    m, ncols = 1200, 37
    if case == "random":
        k0, k1 = 100, 356
        ipiv = np.arange(m)
        ipiv[k0:k1] = [rng.integers(s, m) for s in range(k0, k1)]
    elif case == "repeat_outside":
        # the same row below the block is hit again and again (rows move back and forth)
        k0, k1 = 0, 200
        ipiv = np.arange(m)
        ipiv[k0:k1] = 900
    elif case == "chain":
        k0, k1 = 10, 266
        ipiv = np.arange(m)
        ipiv[k0:k1] = [min(s + 1 + (s % 3), m - 1) for s in range(k0, k1)]
    elif case == "identity":
        k0, k1 = 0, 256
        ipiv = np.arange(m)
    else:
        k0, k1 = 0, 1100  # several 256-pivot blocks
        ipiv = np.array([rng.integers(s, m) for s in range(m)])
    A = rng.uniform(-1, 1, (m, ncols))
    want = seq_swaps(A, ipiv, k0, k1)
    dA = torch.tensor(A.T.copy(), device=dev)  # column-major storage
    dpiv = torch.tensor(ipiv.astype(np.int32), device=dev)
    st = LIB.al_dev_dlaswp(ncols, dA.data_ptr(), m, k0, k1, dpiv.data_ptr(), None)
    assert st == 0, _lib.last_error()
    torch.cuda.synchronize()
    assert np.array_equal(dA.cpu().numpy().T, want)


def test_device_getrf_getrs_on_a_caller_stream():
    dev = _dev()
    n, nrhs = 700, 3
    ab = O.larnv(O.SYMMETRIC, [1, 2, 3, 5], n * n + n * nrhs)[0]
    a0, b0 = ab[: n * n], ab[n * n:]
    s = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(s):
        A = torch.tensor(a0, device=dev)
        B = torch.tensor(b0, device=dev)
        ipiv = torch.empty(n, dtype=torch.int32, device=dev)
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        sp = ctypes.c_void_p(s.cuda_stream)
        assert LIB.al_dev_dgetrf(n, n, A.data_ptr(), n, ipiv.data_ptr(), info.data_ptr(), sp) == 0
        assert LIB.al_dev_dgetrs(n, nrhs, A.data_ptr(), n, ipiv.data_ptr(), B.data_ptr(), n, sp) == 0
    s.synchronize()
    assert int(info.item()) == 0
    x = B.cpu().numpy()
    assert O.backward_error(a0, x, b0, n, nrhs) <= 1e-14
    lu = A.cpu().numpy()
    st, lu_o, piv_o, _ = O.dgetrf(a0, n)
    assert O.get01(a0, lu, ipiv.cpu().numpy(), n, n) < 30


def test_profiler_counts_launches():
    dev = _dev()
    m = n = k = 256
    A = torch.rand(m * k, dtype=torch.float64, device=dev)
    C = torch.zeros(m * n, dtype=torch.float64, device=dev)
    LIB.al_prof_reset()
    LIB.al_prof_enable(1)
    for _ in range(3):
        assert LIB.al_dev_dgemm(0, 0, m, n, k, 1.0, A.data_ptr(), m, A.data_ptr(), k, 0.0, C.data_ptr(), m, None) == 0
    LIB.al_prof_enable(0)
    L, ms, fl, by = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    assert LIB.al_prof_query(b"dgemm_nn", ctypes.byref(L), ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(by)) == 0
    LIB.al_prof_reset()
    assert L.value == 3 and ms.value > 0 and fl.value == 3 * 2.0 * m * n * k


# ------------------------------------------------------------------------------------------
# low-precision GEMMs (the building blocks of the mixed-precision factorization)
# ------------------------------------------------------------------------------------------

def bf16_round(x32: np.ndarray) -> np.ndarray:
    """float32 -> nearest-even bfloat16, returned as float32 (numpy restatement)."""
    b = np.ascontiguousarray(x32, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def _f32_dev(a2d: np.ndarray, ld: int, dev):
    """Column-major FP32 device copy of a 2-D array with leading dimension ld."""
    m, n = a2d.shape
    buf = np.zeros(ld * max(n, 1), dtype=np.float32)
    for j in range(n):
        buf[j * ld:j * ld + m] = a2d[:, j]
    return torch.from_numpy(buf).to(dev)


def _from_dev(t, m, n, ld):
    h = t.cpu().numpy()
    return np.stack([h[j * ld:j * ld + m] for j in range(n)], axis=1) if n else np.zeros((m, 0), np.float32)


def test_bf16_rounding_restatement_is_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5], dtype=np.float32)
    got = bf16_round(x)
    assert got[0] == 1.0
    assert got[1] == 1.0            # tie between 1 and 1 + 2^-7: to even
    assert got[2] == 1.0 + 2.0 ** -6  # tie: to even (upwards)
    assert got[3] == 1.0
    assert got[4] == -2.5
    assert np.array_equal(got, torch.from_numpy(x).to(torch.bfloat16).float().numpy())


@pytest.mark.parametrize("m,n,k", [(130, 70, 100), (256, 256, 512), (1000, 513, 300), (64, 64, 64), (5, 3, 1)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.0, 1.0), (0.75, -0.5)])
def test_gemm_bf16_equals_fp32_accumulation_of_rounded_operands(m, n, k, alpha, beta):
    dev = _dev()
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    B = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    C0 = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    lda, ldb, ldc = m + 3, k + 1, m + 5
    dA, dB, dC = _f32_dev(A, lda, dev), _f32_dev(B, ldb, dev), _f32_dev(C0, ldc, dev)
    rc = LIB.al_dev_gemm_bf16(m, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta, dC.data_ptr(), ldc, None)
    assert rc == 0, _lib.last_error()
    torch.cuda.synchronize()
    got = _from_dev(dC, m, n, ldc).astype(np.float64)
    Ab, Bb = bf16_round(A).astype(np.float64), bf16_round(B).astype(np.float64)
    exact = alpha * (Ab @ Bb) + beta * C0.astype(np.float64)
    # FP32 accumulation of exact bf16 x bf16 products, then the FP32 alpha/beta epilogue
    bound = 2 * (k + 2) * 2.0 ** -24 * (abs(alpha) * (np.abs(Ab) @ np.abs(Bb)) + abs(beta) * np.abs(C0))
    assert np.all(np.abs(got - exact) <= bound + 1e-30)
    # and it is NOT the FP32 product when the rounding matters (operands really are bf16)
    if k >= 64:
        full = alpha * (A.astype(np.float64) @ B.astype(np.float64)) + beta * C0
        assert np.max(np.abs(got - full)) > np.max(bound)


@pytest.mark.parametrize("ta", [0, 1])
@pytest.mark.parametrize("tb", [0, 1])
@pytest.mark.parametrize("m,n,k", [(130, 70, 100), (512, 256, 384), (7, 300, 65)])
def test_sgemm_fp32_matrix_cores(ta, tb, m, n, k):
    dev = _dev()
    rng = np.random.default_rng(m + n + k + 10 * ta + 20 * tb)
    Aop = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    Bop = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    C0 = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    A = Aop.T.copy() if ta else Aop
    B = Bop.T.copy() if tb else Bop
    lda, ldb, ldc = A.shape[0] + 2, B.shape[0] + 1, m
    dA, dB, dC = _f32_dev(A, lda, dev), _f32_dev(B, ldb, dev), _f32_dev(C0, ldc, dev)
    rc = LIB.al_dev_sgemm(ta, tb, m, n, k, 1.5, dA.data_ptr(), lda, dB.data_ptr(), ldb, -0.5, dC.data_ptr(), ldc, None)
    assert rc == 0, _lib.last_error()
    torch.cuda.synchronize()
    got = _from_dev(dC, m, n, ldc).astype(np.float64)
    exact = 1.5 * (Aop.astype(np.float64) @ Bop.astype(np.float64)) - 0.5 * C0
    bound = 2 * (k + 2) * 2.0 ** -24 * (1.5 * (np.abs(Aop).astype(np.float64) @ np.abs(Bop)) + 0.5 * np.abs(C0))
    assert np.all(np.abs(got - exact) <= bound)


@pytest.mark.parametrize("beta", [0.0, -0.5])
@pytest.mark.parametrize("m,n,k,pad", [(132, 70, 100, 4), (512, 257, 384, 0), (4, 1, 4, 8), (260, 130, 36, 12),
                                       (16320, 64, 64, 64)])
def test_sgemm_untransposed_16_byte_path(m, n, k, pad, beta):
    # shapes with m % 4 == 0, k % 4 == 0 and leading dimensions that are multiples of 4 take the
    # float4 kernel; edge tiles in every dimension and K not a multiple of its 32-step
    dev = _dev()
    rng = np.random.default_rng(m + 3 * n + 7 * k)
    A = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    B = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    C0 = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    lda, ldb, ldc = m + pad, k + pad, m + pad
    dA, dB, dC = _f32_dev(A, lda, dev), _f32_dev(B, ldb, dev), _f32_dev(C0, ldc, dev)
    c_before = dC.cpu().numpy()
    rc = LIB.al_dev_sgemm(0, 0, m, n, k, 1.5, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta, dC.data_ptr(), ldc, None)
    assert rc == 0, _lib.last_error()
    torch.cuda.synchronize()
    got = _from_dev(dC, m, n, ldc).astype(np.float64)
    exact = 1.5 * (A.astype(np.float64) @ B.astype(np.float64)) + beta * C0
    bound = 2 * (k + 2) * 2.0 ** -24 * (1.5 * (np.abs(A).astype(np.float64) @ np.abs(B)) + abs(beta) * np.abs(C0))
    assert np.all(np.abs(got - exact) <= bound)
    # the padding rows between columns are untouched
    after = dC.cpu().numpy()
    mask = np.ones_like(after, dtype=bool)
    for j in range(n):
        mask[j * ldc:j * ldc + m] = False
    np.testing.assert_array_equal(after[mask], c_before[mask])
