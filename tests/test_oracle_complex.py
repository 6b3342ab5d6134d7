"""Pins the complex twins of the CPU oracle (or_zgemm, or_zgetrf, or_zgetrf_rect, or_zgesv,
or_zpotrf in oracle/al_oracle.c) before any GPU result is checked against them.

The checks, chosen so that a plausible transcription mistake fails one of them:
  * the crate's CBLAS-captured complex products (src/blas.rs:2029-2195), at the crate's own
    tolerance 1e-13 relative to max(|want|, 1) (src/blas.rs:2386-2389);
  * numpy's complex matmul on explicitly transposed operands, over the crate's shape sweep
    (src/blas.rs:2320-2342), with and without beta's NaN;
  * scipy's LAPACK zgetrf (izamax pivots on |Re| + |Im|), zpotrf and numpy's solve;
  * the crate's own test matrices and answers (src/lapack/complex.rs:422-646);
  * planted mistakes (a conjugated transpose, a modulus pivot, a missing conjugate in the
    Cholesky update) that these checks must reject.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
ZL = FIX["zlapack"]


def cx(pairs) -> np.ndarray:
    return np.array([complex(a, b) for a, b in pairs], dtype=np.complex128)


def cm(flat: np.ndarray, m: int, n: int) -> np.ndarray:
    """Column-major flat slice -> (m, n) matrix."""
    return np.asarray(flat).reshape(n, m).T


def flat(M: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(M).T).ravel()


def close_cpx(got, want, tol=1e-13) -> bool:
    got, want = np.asarray(got), np.asarray(want)
    scale = np.maximum(np.abs(want), 1.0)
    return bool(np.all(np.abs(got - want) / scale < tol))


# ------------------------------------------------------------------------------------------
# zgemm family
# ------------------------------------------------------------------------------------------

ZGEMM_FLAGS = {"zgemm_tn": (True, False), "zgemm_nt": (False, True), "zgemm_tt": (True, True)}


@pytest.mark.parametrize("name", sorted(ZGEMM_FLAGS))
def test_zgemm_variants_match_the_cblas_fixtures(name):
    f = FIX[name]
    ta, tb = ZGEMM_FLAGS[name]
    got = O.zgemm(ta, tb, complex(*f["alpha"]), cx(f["a"]), cx(f["b"]), complex(*f["beta"]), cx(f["c0"]),
                  f["m"], f["n"], f["k"])
    assert close_cpx(got, cx(f["want"])), (got, f["want"])


def _zgemm_conjugating(ta, tb, alpha, a, b, beta, c, m, n, k):
    # planted mistake: the transposed operands conjugated ('C' instead of 'T')
    A = cm(a, k, m).conj().T if ta else cm(a, m, k)
    B = cm(b, n, k).conj().T if tb else cm(b, k, n)
    return flat(alpha * (A @ B) + beta * cm(c, m, n))


def test_zgemm_fixtures_reject_a_conjugated_transpose():
    for name, (ta, tb) in ZGEMM_FLAGS.items():
        f = FIX[name]
        bad = _zgemm_conjugating(ta, tb, complex(*f["alpha"]), cx(f["a"]), cx(f["b"]), complex(*f["beta"]),
                                 cx(f["c0"]), f["m"], f["n"], f["k"])
        assert not close_cpx(bad, cx(f["want"])), name


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (1, 4, 1), (3, 2, 5), (2, 3, 4), (5, 5, 5), (7, 11, 13)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_zgemm_matches_the_definition_on_the_crate_sweep(m, n, k, ta, tb):
    # the crate's sweep data (src/blas.rs:2326-2340)
    a_op = np.array([complex((i % 3) - 1.0, (i % 5) - 2.0) for i in range(m * k)])
    b_op = np.array([complex((i % 4) - 1.5, (i % 7) - 3.0) for i in range(k * n)])
    c0 = np.array([complex(i % 2, 0.5) for i in range(m * n)])
    alpha, beta = complex(0.5, -1.5), complex(2.0, 0.25)
    A, B = cm(a_op, m, k), cm(b_op, k, n)
    a_st = flat(A.T) if ta else a_op
    b_st = flat(B.T) if tb else b_op
    got = O.zgemm(ta, tb, alpha, a_st, b_st, beta, c0, m, n, k)
    want = flat(alpha * (A @ B) + beta * cm(c0, m, n))
    # small integers and halves: every product and partial sum is exact
    assert np.array_equal(got, want)


def test_zgemm_beta_times_c_is_always_evaluated():
    # c[idx] = alpha*sum + beta*c[idx] (src/blas.rs:1115-1116): a NaN in C survives beta = 0,
    # and k = 0 leaves beta*C
    a = np.array([1 + 1j, 2 - 1j])
    b = np.array([3 + 0j, 1 + 2j])
    c = np.array([np.nan + 0j])
    got = O.zgemm(False, False, 1.0, a, b, 0.0, c, 1, 1, 2)
    assert np.isnan(got[0].real)
    got = O.zgemm(False, False, 2.0, np.zeros(0), np.zeros(0), 1j, np.array([3 + 4j]), 1, 1, 0)
    assert got[0] == 1j * (3 + 4j)


def test_zgemm_random_against_numpy():
    rng = np.random.default_rng(7)
    m, n, k = 19, 23, 31
    A = rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))
    B = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
    C = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    alpha, beta = 0.75 - 0.5j, -1.25 + 2j
    got = O.zgemm(False, True, alpha, flat(A), flat(B.T), beta, flat(C), m, n, k)
    assert close_cpx(got, flat(alpha * (A @ B) + beta * C), 1e-13)


# ------------------------------------------------------------------------------------------
# zgetrf / zgetrf_rect / zgesv
# ------------------------------------------------------------------------------------------

def unpack_lu(lu, m, n):
    LU = cm(lu, m, n)
    kk = min(m, n)
    L = np.tril(LU[:, :kk], -1) + np.eye(m, kk)
    U = np.triu(LU[:kk, :])
    return L, U


def apply_pivots(A, ipiv):
    PA = A.copy()
    for k, p in enumerate(ipiv):
        PA[[k, p], :] = PA[[p, k], :]
    return PA


def test_zgetrf_reconstructs_general3_and_the_last_pivot_is_n_minus_1():
    a = cx(ZL["general3"])
    st, lu, piv = O.zgetrf(a, 3)
    assert st == 0 and piv[2] == 2
    L, U = unpack_lu(lu, 3, 3)
    assert np.max(np.abs(L @ U - apply_pivots(cm(a, 3, 3), piv))) < 1e-12


def test_zgetrf_reports_the_crate_singular_matrix():
    st, _, _ = O.zgetrf(cx(ZL["singular2"]), 2)
    assert st == 1


def _pivot_by_modulus(col):
    return int(np.argmax(np.abs(col)))


def test_zgetrf_pivots_on_abs1_not_the_modulus():
    # |4.5| = 4.5 > |3 + 2.5i| = 3.9, but |Re| + |Im| is 4.5 < 5.5: izamax takes row 1
    a = np.array([4.5 + 0j, 3 + 2.5j, 1 + 0j, 2 + 0j])
    st, _, piv = O.zgetrf(a, 2)
    assert st == 0 and piv[0] == 1
    assert _pivot_by_modulus(a[:2]) != piv[0]  # the planted mistake picks another row


@pytest.mark.parametrize("n,seed", [(8, 1), (33, 2), (70, 3)])
def test_zgetrf_matches_scipy_lapack(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    st, lu, piv = O.zgetrf(flat(A), n)
    assert st == 0
    lu_s, piv_s = sla.lu_factor(A)
    assert np.array_equal(piv, piv_s)
    assert close_cpx(lu, flat(lu_s), 1e-11)


def test_zgetrf_rect_reconstructs_the_crate_panel_and_equals_zgetrf_when_square():
    nr, nc = ZL["rect_shape"]
    a0 = cx(ZL["rect3x2"])
    st, lu, piv = O.zgetrf_rect(a0, nr, nc)
    assert st == 0
    L, U = unpack_lu(lu, nr, nc)
    assert np.max(np.abs(L @ U - apply_pivots(cm(a0, nr, nc), piv))) < 1e-12
    # the square case performs the same operations on every element, in the same order
    rng = np.random.default_rng(11)
    A = rng.standard_normal((12, 12)) + 1j * rng.standard_normal((12, 12))
    s1, lu1, p1 = O.zgetrf(flat(A), 12)
    s2, lu2, p2 = O.zgetrf_rect(flat(A), 12, 12)
    assert s1 == s2 == 0 and np.array_equal(p1, p2) and np.array_equal(lu1, lu2)


def test_zgetrf_rect_matches_scipy_on_a_tall_panel_and_rejects_wide():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((90, 17)) + 1j * rng.standard_normal((90, 17))
    st, lu, piv = O.zgetrf_rect(flat(A), 90, 17)
    assert st == 0
    lu_s, piv_s = sla.lu_factor(A)
    assert np.array_equal(piv, piv_s)
    assert close_cpx(lu, flat(lu_s), 1e-11)
    st, _, _ = O.zgetrf_rect(np.ones(2, dtype=complex), 1, 2)
    assert st == 2


def test_zgesv_recovers_the_crate_known_solution():
    a = cx(ZL["general3"])
    x_true = cx(ZL["zgesv_x_true"])
    b = cm(a, 3, 3) @ x_true
    st, x = O.zgesv(a, b)
    assert st == 0 and np.max(np.abs(x - x_true)) < 1e-12
    st, _ = O.zgesv(cx(ZL["singular2"]), np.ones(2, dtype=complex))
    assert st == 1


def test_zgesv_matches_numpy_solve():
    rng = np.random.default_rng(3)
    n = 60
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.eye(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    st, x = O.zgesv(flat(A), b)
    assert st == 0
    assert close_cpx(x, np.linalg.solve(A, b), 1e-12)


# ------------------------------------------------------------------------------------------
# zpotrf
# ------------------------------------------------------------------------------------------

def _zpotrf_lower_without_conj(A):
    # planted mistake: the update uses a_ik * a_jk instead of a_ik * conj(a_jk)
    n = A.shape[0]
    L = np.tril(A).astype(complex)
    for j in range(n):
        d = (L[j, j].real - sum(abs(L[j, k]) ** 2 for k in range(j))) ** 0.5
        L[j, j] = d
        for i in range(j + 1, n):
            L[i, j] = (L[i, j] - sum(L[i, k] * L[j, k] for k in range(j))) / d
    return L


@pytest.mark.parametrize("upper", [False, True])
def test_zpotrf_factorises_spd3_and_leaves_the_other_triangle(upper):
    a = cx(ZL["spd3"])
    A = cm(a, 3, 3)
    poisoned = A.copy()
    other = np.tril_indices(3, -1) if upper else np.triu_indices(3, 1)
    poisoned[other] = np.nan
    st, f = O.zpotrf(flat(poisoned), 3, upper)
    assert st == 0
    F = cm(f, 3, 3)
    assert np.all(np.isnan(F[other]))  # untouched
    if upper:
        U = np.triu(F)
        assert np.max(np.abs(U.conj().T @ U - A)) < 1e-12
    else:
        L = np.tril(F)
        assert np.max(np.abs(L @ L.conj().T - A)) < 1e-12
    # the check rejects a factor whose update drops the conjugate
    Lbad = _zpotrf_lower_without_conj(A)
    assert np.max(np.abs(Lbad @ Lbad.conj().T - A)) > 1e-6


def test_zpotrf_rejects_the_crate_indefinite_matrix_and_writes_the_pivot():
    # [[1, 2], [2, 1]]: column 0 gives l11 = 1, l21 = 2; the next pivot is 1 - |2|^2 = -3,
    # which is written to the diagonal (imaginary part 0) before Singular is returned
    st, f = O.zpotrf(cx(ZL["not_pd2"]), 2, False)
    assert st == 1 and f[3] == complex(-3.0, 0.0) and f[0] == 1.0 and f[1] == 2.0
    st, f = O.zpotrf(cx(ZL["not_pd2"]), 2, True)
    assert st == 1 and f[3] == complex(-3.0, 0.0) and f[2] == 2.0


def test_zpotrf_nan_diagonal_is_singular():
    a = np.array([np.nan + 0j, 0j, 0j, 1 + 0j])
    st, f = O.zpotrf(a, 2, False)
    assert st == 1 and np.isnan(f[0].real)


@pytest.mark.parametrize("upper", [False, True])
def test_zpotrf_matches_scipy_lapack(upper):
    rng = np.random.default_rng(21)
    n = 40
    G = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = G @ G.conj().T + n * np.eye(n)
    st, f = O.zpotrf(flat(A), n, upper)
    assert st == 0
    F = cm(f, n, n)
    want = sla.cholesky(A, lower=not upper)
    got = np.triu(F) if upper else np.tril(F)
    assert close_cpx(flat(got), flat(want), 1e-12)


def test_empty_complex_cases():
    st, lu, piv = O.zgetrf(np.zeros(0, dtype=complex), 0)
    assert st == 0 and piv.size == 0
    st, x = O.zgesv(np.zeros(0, dtype=complex), np.zeros(0, dtype=complex))
    assert st == 0 and x.size == 0
    st, _ = O.zpotrf(np.zeros(0, dtype=complex), 0, False)
    assert st == 0
