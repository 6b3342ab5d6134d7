"""Complex twins on the GPU (zgemm*, zgetrf, zgetrf_rect, zgesv, zpotrf) against the oracle.

Bars: bit for bit where the GPU path performs the crate's own operations in the crate's order
(an LU whose columns fit one 32-wide complex panel, any number of rows; a Cholesky of order
<= 64); elsewhere the crate's test tolerance 1e-13 relative to max(|want|, 1) on its fixtures,
an elementwise rounding bound for products, and LAPACK's normalized residuals for
factorizations (threshold 30, as TESTING/LIN uses).  torch is used for device memory and, at
sizes the oracle cannot reach in seconds, for an independent complex128 matmul on the GPU.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
al = pytest.importorskip("algebra_linear_amd")
from algebra_linear_amd import _lib  # noqa: E402
from algebra_linear_amd.errors import LapackError  # noqa: E402

LIB = _lib.load()
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
ZL = FIX["zlapack"]
EPS = np.finfo(float).eps


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def cx(pairs):
    return np.array([complex(a, b) for a, b in pairs], dtype=np.complex128)


def cm(flat, m, n):
    return np.asarray(flat).reshape(n, m).T


def flat(M):
    return np.ascontiguousarray(np.asarray(M).T).ravel()


def crand(rng, *shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def close_cpx(got, want, tol=1e-13):
    scale = np.maximum(np.abs(want), 1.0)
    return bool(np.all(np.abs(np.asarray(got) - np.asarray(want)) / scale < tol))


def lu_residual(A, lu, ipiv):
    """||P A - L U||_1 / (n ||A||_1 eps), DGET01's ratio for a complex factorization."""
    m, n = A.shape
    kk = min(m, n)
    LU = cm(lu, m, n)
    L = np.tril(LU[:, :kk], -1) + np.eye(m, kk)
    U = np.triu(LU[:kk, :])
    PA = A.copy()
    for k, p in enumerate(ipiv):
        k, p = int(k), int(p)
        PA[[k, p], :] = PA[[p, k], :]
    return np.abs(PA - L @ U).sum(axis=0).max() / (n * np.abs(A).sum(axis=0).max() * EPS)


ZGEMM = {"zgemm": (al.blas.zgemm, False, False), "zgemm_tn": (al.blas.zgemm_tn, True, False),
         "zgemm_nt": (al.blas.zgemm_nt, False, True), "zgemm_tt": (al.blas.zgemm_tt, True, True)}


# ------------------------------------------------------------------------------------------
# zgemm
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["zgemm_tn", "zgemm_nt", "zgemm_tt"])
def test_zgemm_matches_the_crate_cblas_fixtures(name):
    _dev()
    f = FIX[name]
    fn, _, _ = ZGEMM[name]
    c = cx(f["c0"])
    fn(complex(*f["alpha"]), cx(f["a"]), cx(f["b"]), complex(*f["beta"]), c, f["m"], f["n"], f["k"])
    assert close_cpx(c, cx(f["want"])), (c, f["want"])


@pytest.mark.parametrize("name", sorted(ZGEMM))
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (1, 4, 1), (3, 2, 5), (7, 11, 13), (129, 65, 33), (300, 257, 190)])
def test_zgemm_matches_the_oracle(name, m, n, k):
    _dev()
    fn, ta, tb = ZGEMM[name]
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A, B, C = crand(rng, m, k), crand(rng, k, n), crand(rng, m, n)
    alpha, beta = 0.5 - 1.5j, 2.0 + 0.25j
    a = flat(A.T) if ta else flat(A)
    b = flat(B.T) if tb else flat(B)
    want = O.zgemm(ta, tb, alpha, a, b, beta, flat(C), m, n, k)
    got = flat(C).copy()
    fn(alpha, a, b, beta, got, m, n, k)
    bound = 8 * (k + 2) * EPS * (abs(alpha) * flat(np.abs(A) @ np.abs(B)) + abs(beta) * np.abs(flat(C))) + 1e-300
    assert np.all(np.abs(got - want) <= bound)


def test_zgemm_forms_beta_times_c_always_and_handles_k_zero():
    _dev()
    c = np.array([np.nan + 0j, 1 + 1j])
    al.blas.zgemm(1.0, np.array([1 + 1j]), np.array([2 + 0j, 0 + 3j]), 0.0, c, 1, 2, 1)
    assert np.isnan(c[0].real) and c[1] == (1 + 1j) * (0 + 3j)
    c = np.array([3 + 4j, -1 + 2j])
    al.blas.zgemm(2.0, np.zeros(0, complex), np.zeros(0, complex), 1j, c, 2, 1, 0)
    assert np.array_equal(c, 1j * np.array([3 + 4j, -1 + 2j]))


def test_zgemm_large_against_torch():
    dev = _dev()
    rng = np.random.default_rng(99)
    m, n, k = 1536, 1280, 1024
    A, B, C = crand(rng, m, k), crand(rng, k, n), crand(rng, m, n)
    alpha, beta = 1.0 - 0.5j, -0.75 + 0.0j
    dA = torch.tensor(A.T.copy(), device=dev)  # column-major storage
    dB = torch.tensor(B.T.copy(), device=dev)
    dC = torch.tensor(C.T.copy(), device=dev)
    st = LIB.al_dev_zgemm(0, 0, m, n, k, _lib.Complex.of(alpha), dA.data_ptr(), m, dB.data_ptr(), k,
                          _lib.Complex.of(beta), dC.data_ptr(), m, None)
    assert st == 0, _lib.last_error()
    torch.cuda.synchronize()
    ref = alpha * (torch.tensor(A, device=dev) @ torch.tensor(B, device=dev)) + beta * torch.tensor(C, device=dev)
    got = dC.T.cpu().numpy()
    bound = 8 * (k + 2) * EPS * (abs(alpha) * (np.abs(A) @ np.abs(B)) + abs(beta) * np.abs(C))
    assert np.all(np.abs(got - ref.cpu().numpy()) <= 2 * bound)


def test_dev_zgemm_on_submatrices_leaves_the_rest():
    dev = _dev()
    rng = np.random.default_rng(3)
    m, n, k, ld = 70, 50, 40, 96
    A, B, C = crand(rng, ld, k), crand(rng, ld, n), crand(rng, ld, n)
    dA = torch.tensor(A.T.copy(), device=dev)
    dB = torch.tensor(B.T.copy(), device=dev)
    dC = torch.tensor(C.T.copy(), device=dev)
    alpha, beta = 0.3 + 0.7j, 1.0 - 1.0j
    st = LIB.al_dev_zgemm(0, 1, m, n, k, _lib.Complex.of(alpha), dA.data_ptr(), ld, dB.data_ptr(), ld,
                          _lib.Complex.of(beta), dC.data_ptr(), ld, None)
    assert st == 0, _lib.last_error()
    got = dC.T.cpu().numpy()
    want = alpha * (A[:m, :k] @ B[:n, :k].T) + beta * C[:m, :n]
    assert close_cpx(got[:m, :n], want, 1e-13)
    assert np.array_equal(got[m:, :], C[m:, :]) and np.array_equal(got[:, n:], C[:, n:])


# ------------------------------------------------------------------------------------------
# zgetrf / zgetrf_rect
# ------------------------------------------------------------------------------------------

def test_zgetrf_crate_fixtures():
    _dev()
    a = cx(ZL["general3"])
    lu = a.copy()
    piv = al.lapack.zgetrf(lu, 3)
    st, want, wpiv = O.zgetrf(a, 3)
    assert st == 0 and piv == list(wpiv) and np.array_equal(lu, want)
    with pytest.raises(LapackError) as e:
        al.lapack.zgetrf(cx(ZL["singular2"]), 2)
    assert e.value.kind == "Singular"
    nr, nc = ZL["rect_shape"]
    a0 = cx(ZL["rect3x2"])
    lu = a0.copy()
    ipiv = np.zeros(nc, dtype=np.uintp)
    al.lapack.zgetrf_rect(lu, nr, nc, ipiv)
    st, want, wpiv = O.zgetrf_rect(a0, nr, nc)
    assert np.array_equal(lu, want) and list(ipiv) == list(wpiv)


@pytest.mark.parametrize("n", [1, 2, 5, 17, 31, 32])
def test_zgetrf_one_panel_is_bit_exact(n):
    _dev()
    rng = np.random.default_rng(n)
    a = flat(crand(rng, n, n))
    lu = a.copy()
    piv = al.lapack.zgetrf(lu, n)
    st, want, wpiv = O.zgetrf(a, n)
    assert st == 0 and piv == list(wpiv)
    assert np.array_equal(lu, want)


@pytest.mark.parametrize("nrows,ncols", [(300, 32), (1000, 17), (3000, 32), (20000, 8)])
def test_zgetrf_rect_tall_one_panel_is_bit_exact(nrows, ncols):
    _dev()
    rng = np.random.default_rng(nrows + ncols)
    a = flat(crand(rng, nrows, ncols))
    lu = a.copy()
    ipiv = np.zeros(ncols, dtype=np.uintp)
    al.lapack.zgetrf_rect(lu, nrows, ncols, ipiv)
    st, want, wpiv = O.zgetrf_rect(a, nrows, ncols)
    assert st == 0 and list(ipiv) == list(wpiv)
    assert np.array_equal(lu, want)


@pytest.mark.parametrize("n", [33, 100, 257, 700])
def test_zgetrf_blocked_against_the_oracle(n):
    _dev()
    rng = np.random.default_rng(7 * n)
    A = crand(rng, n, n)
    a = flat(A)
    lu = a.copy()
    piv = al.lapack.zgetrf(lu, n)
    st, want, wpiv = O.zgetrf(a, n)
    assert st == 0
    assert lu_residual(A, lu, piv) < 30
    agree = np.mean(np.array(piv) == wpiv)
    assert agree >= 0.95, agree
    if agree == 1.0:
        assert close_cpx(lu, want, 1e-9)


@pytest.mark.parametrize("nrows,ncols", [(2000, 100), (5000, 257)])
def test_zgetrf_rect_blocked(nrows, ncols):
    _dev()
    rng = np.random.default_rng(nrows)
    A = crand(rng, nrows, ncols)
    lu = flat(A).copy()
    ipiv = np.zeros(ncols, dtype=np.uintp)
    al.lapack.zgetrf_rect(lu, nrows, ncols, ipiv)
    assert lu_residual(A, lu, list(ipiv)) < 30


def test_zgetrf_singular_in_a_later_panel():
    _dev()
    rng = np.random.default_rng(8)
    n = 150
    A = crand(rng, n, n)
    A[:, 90] = 0  # every update of column 90 subtracts l * 0: its pivot at step 90 is exactly zero
    st, _, _ = O.zgetrf(flat(A), n)
    assert st == 1
    with pytest.raises(LapackError) as e:
        al.lapack.zgetrf(flat(A), n)
    assert e.value.kind == "Singular"
    dev = _dev()
    dA = torch.tensor(A.T.copy(), device=dev)
    piv = torch.zeros(n, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    assert LIB.al_dev_zgetrf(n, n, dA.data_ptr(), n, piv.data_ptr(), info.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert int(info.item()) == 91


def test_zgetrf_nan_rules_match_the_oracle():
    _dev()
    rng = np.random.default_rng(12)
    for case in ("diag", "below"):
        n = 20
        A = crand(rng, n, n)
        if case == "diag":
            A[0, 0] = complex(np.nan, 0.0)  # a NaN start blocks every later candidate
        else:
            A[5, 0] = complex(0.0, np.nan)  # a later NaN is skipped
        a = flat(A)
        lu = a.copy()
        piv = al.lapack.zgetrf(lu, n)
        st, want, wpiv = O.zgetrf(a, n)
        assert st == 0 and piv == list(wpiv)
        assert np.array_equal(np.isnan(lu), np.isnan(want))
        ok = ~np.isnan(want)
        assert np.array_equal(lu[ok], want[ok])


# ------------------------------------------------------------------------------------------
# zgesv and the device solve
# ------------------------------------------------------------------------------------------

def test_zgesv_crate_known_solution_and_singular():
    _dev()
    a = cx(ZL["general3"])
    x_true = cx(ZL["zgesv_x_true"])
    b = cm(a, 3, 3) @ x_true
    aa = a.copy()
    al.lapack.zgesv(aa, b)
    assert np.max(np.abs(b - x_true)) < 1e-12
    bs = np.ones(2, dtype=complex)
    with pytest.raises(LapackError) as e:
        al.lapack.zgesv(cx(ZL["singular2"]), bs)
    assert e.value.kind == "Singular" and np.array_equal(bs, np.ones(2, dtype=complex))


@pytest.mark.parametrize("n", [1, 32, 200, 1000])
def test_zgesv_against_the_oracle(n):
    _dev()
    rng = np.random.default_rng(n + 5)
    A = crand(rng, n, n) + 0.5 * np.eye(n)
    b = crand(rng, n)
    x = b.copy()
    al.lapack.zgesv(flat(A).copy(), x)
    st, want = O.zgesv(flat(A), b)
    assert st == 0
    r = b - A @ x
    be = np.abs(r).max() / (np.abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())
    assert be < 1e-14, be
    cond = np.linalg.cond(A, 1)
    assert np.abs(x - want).max() / np.abs(want).max() < 100 * n * EPS * cond


@pytest.mark.parametrize("nrhs", [3, 20])
def test_dev_zgetrf_zgetrs_multi_rhs(nrhs):
    dev = _dev()
    rng = np.random.default_rng(nrhs)
    n = 600
    A = crand(rng, n, n)
    X = crand(rng, n, nrhs)
    B = A @ X
    dA = torch.tensor(A.T.copy(), device=dev)
    dB = torch.tensor(B.T.copy(), device=dev)
    piv = torch.zeros(n, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    assert LIB.al_dev_zgetrf(n, n, dA.data_ptr(), n, piv.data_ptr(), info.data_ptr(), None) == 0
    assert LIB.al_dev_zgetrs(n, nrhs, dA.data_ptr(), n, piv.data_ptr(), dB.data_ptr(), n, None) == 0
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    got = dB.T.cpu().numpy()
    cond = np.linalg.cond(A, 1)
    assert np.abs(got - X).max() / np.abs(X).max() < 100 * n * EPS * cond


def test_zgesv_large_backward_error():
    dev = _dev()
    rng = np.random.default_rng(2024)
    n = 4096
    A = crand(rng, n, n)
    b = crand(rng, n)
    dA = torch.tensor(A.T.copy(), device=dev)
    dB = torch.tensor(b, device=dev)
    piv = torch.zeros(n, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    assert LIB.al_dev_zgetrf(n, n, dA.data_ptr(), n, piv.data_ptr(), info.data_ptr(), None) == 0
    assert LIB.al_dev_zgetrs(n, 1, dA.data_ptr(), n, piv.data_ptr(), dB.data_ptr(), n, None) == 0
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    x = dB.cpu().numpy()
    At = torch.tensor(A, device=dev)
    r = (torch.tensor(b, device=dev) - At @ torch.tensor(x, device=dev)).abs().max().item()
    be = r / (np.abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())
    assert be < 1e-14, be


# ------------------------------------------------------------------------------------------
# zpotrf
# ------------------------------------------------------------------------------------------

def hpd(rng, n):
    G = crand(rng, n, n)
    return G @ G.conj().T + n * np.eye(n)


@pytest.mark.parametrize("upper", [False, True])
def test_zpotrf_crate_fixtures(upper):
    _dev()
    a = cx(ZL["spd3"])
    f = a.copy()
    al.lapack.zpotrf(f, 3, upper)
    st, want = O.zpotrf(a, 3, upper)
    assert st == 0 and np.array_equal(f, want)
    bad = cx(ZL["not_pd2"])
    with pytest.raises(LapackError) as e:
        al.lapack.zpotrf(bad, 2, upper)
    assert e.value.kind == "Singular" and bad[3] == complex(-3.0, 0.0)


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("n", [1, 7, 33, 64])
def test_zpotrf_one_block_is_bit_exact(n, upper):
    _dev()
    rng = np.random.default_rng(n + (100 if upper else 0))
    A = hpd(rng, n)
    poisoned = A.copy()
    other = np.tril_indices(n, -1) if upper else np.triu_indices(n, 1)
    poisoned[other] = np.nan
    f = flat(poisoned).copy()
    al.lapack.zpotrf(f, n, upper)
    st, want = O.zpotrf(flat(poisoned), n, upper)
    assert st == 0
    assert np.array_equal(np.isnan(f), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.array_equal(f[ok], want[ok])


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("n", [65, 300, 700])
def test_zpotrf_blocked_against_the_oracle(n, upper):
    _dev()
    rng = np.random.default_rng(3 * n + upper)
    A = hpd(rng, n)
    poisoned = A.copy()
    other = np.tril_indices(n, -1) if upper else np.triu_indices(n, 1)
    poisoned[other] = np.nan
    f = flat(poisoned).copy()
    al.lapack.zpotrf(f, n, upper)
    st, want = O.zpotrf(flat(A), n, upper)
    assert st == 0
    F = cm(f, n, n)
    assert np.all(np.isnan(F[other]))  # the other triangle is neither read nor written
    T = np.triu(np.nan_to_num(F)) if upper else np.tril(np.nan_to_num(F))
    R = T.conj().T @ T if upper else T @ T.conj().T
    ratio = np.abs(R - A).sum(axis=0).max() / (n * np.abs(A).sum(axis=0).max() * EPS)
    assert ratio < 30, ratio
    W = np.triu(cm(want, n, n)) if upper else np.tril(cm(want, n, n))
    assert close_cpx(flat(T), flat(W), 1e-10)


@pytest.mark.parametrize("upper", [False, True])
def test_zpotrf_not_positive_definite_fails_at_the_oracle_column(upper):
    dev = _dev()
    rng = np.random.default_rng(31)
    n = 300
    A = hpd(rng, n)
    A[200, 200] = -1.0e6  # the pivot of column 200 goes negative
    st, want = O.zpotrf(flat(A), n, upper)
    assert st == 1
    Wm = cm(want, n, n)
    col = next(j for j in range(n) if Wm[j, j].real <= 0)
    dA = torch.tensor(A.T.copy(), device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    assert LIB.al_dev_zpotrf(int(upper), n, dA.data_ptr(), n, info.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert int(info.item()) == col + 1
    got = dA.T.cpu().numpy()
    assert got[col, col].imag == 0.0
    assert abs(got[col, col].real - Wm[col, col].real) <= 1e-9 * abs(Wm[col, col].real)


def test_zpotrf_large_residual_on_the_gpu():
    dev = _dev()
    n = 2048
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    G = torch.randn(n, n, dtype=torch.complex128, device=dev, generator=g)
    A = G @ G.conj().T + n * torch.eye(n, dtype=torch.complex128, device=dev)
    for upper in (False, True):
        dA = A.T.contiguous()  # column-major storage of A
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        assert LIB.al_dev_zpotrf(int(upper), n, dA.data_ptr(), n, info.data_ptr(), None) == 0
        torch.cuda.synchronize()
        assert int(info.item()) == 0
        F = dA.T
        T = torch.triu(F) if upper else torch.tril(F)
        R = T.conj().T @ T if upper else T @ T.conj().T
        ratio = ((R - A).abs().sum(dim=0).max() / (n * A.abs().sum(dim=0).max() * EPS)).item()
        assert ratio < 30, (upper, ratio)
