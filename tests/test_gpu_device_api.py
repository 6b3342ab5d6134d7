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
