"""Pins the CPU oracle (oracle/al_oracle.c) before anything is checked against it.

Each check is chosen so that a plausible transcription mistake -- a dropped term, a wrong
index, a transposed operand, a reversed swap order -- fails it:
  * golden vectors the crate's own tests hold (tests/golden/reference_fixtures.json);
  * closed forms and hand-computed answers of the crate's tests;
  * definitions computed independently (numpy on an explicitly transposed operand);
  * scipy's LAPACK (OpenBLAS) on the src/parity.rs sample matrices;
  * planted mistakes that the checkers must flag.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O
from matrices import (SWEEP_ALPHA_BETA, SWEEP_SHAPES, cm, hilbert, lower3, mat, sample, sample4, singular3, spd,
                      sweep_operands, transpose_flat, upper3)

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
SEEDY = [1988, 1989, 1990, 1991]
DIST = {"Uniform": O.UNIFORM, "Symmetric": O.SYMMETRIC, "Normal": O.NORMAL}


# ------------------------------------------------------------------------------------------
# generators: the LAPACK stream bit for bit (src/lapack/families.rs:1182-1258)
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("case", FIX["larnv"]["cases"], ids=lambda c: c["dist"])
def test_larnv_matches_the_reference_stream(case):
    x, seed = O.larnv(DIST[case["dist"]], SEEDY, 200)
    for i, w in enumerate(case["first12"]):
        assert x[i] == w, (i, x[i], w)  # bitwise (float equality on finite values)
    assert x[198] == case["last2"][0] and x[199] == case["last2"][1]
    assert list(seed) == case["seed_after"]


@pytest.mark.parametrize("n", [1, 63, 64, 65, 129, 200])
def test_larnv_is_continuous_across_batches(n):
    # src/lapack/families.rs:1261-1278: a batch of n equals n one-at-a-time draws with the
    # seed carried over, and both leave the same seed.
    whole, s_all = O.larnv(O.UNIFORM, SEEDY, n)
    seed = np.array(SEEDY, dtype=np.uint32)
    one = []
    for _ in range(n):
        v, seed = O.larnv(O.UNIFORM, seed, 1)
        one.append(v[0])
    assert np.array_equal(whole, np.array(one))
    assert list(s_all) == list(seed)


@pytest.mark.parametrize("idx", range(12))
def test_dlatms_matches_the_reference_matrices(idx):
    c = FIX["dlatms"]["cases"][idx]
    sym = {"NonSymmetric": O.NONSYMMETRIC, "PositiveSymmetric": O.POSITIVE_SYMMETRIC,
           "Symmetric": O.SYMMETRIC_SYM}[c["sym"]]
    pack = {"Full": O.FULL, "Upper": O.UPPER, "Lower": O.LOWER}[c["pack"]]
    info, a, d, seed = O.dlatms(DIST[c["dist"]], sym, c["mode"], c["cond"], c["dmax"], pack, c["m"], c["n"],
                                SEEDY, d=c["d"])
    assert info == 0
    assert list(seed) == c["seed_after"], "the draws' order and count"
    for got, want in zip(d, c["d"]):
        assert (got == 0.0) if want == 0.0 else abs(got - want) / abs(want) <= 1e-13
    for got, want in zip(a, c["a"]):
        assert (got == 0.0) if want == 0.0 else abs(got - want) / abs(want) <= 1e-13


def test_dlatms_mode3_spectrum_is_geometric_and_matrix_has_it():
    # config-3 generator: sigma_i = cond^(-i/(n-1)); the singular values of U D V are D.
    n, cond = 24, 1e5
    info, a, d, _ = O.dlatms(O.NORMAL, O.NONSYMMETRIC, 3, cond, 1.0, O.FULL, n, n, SEEDY)
    assert info == 0
    want = cond ** (-np.arange(n) / (n - 1.0))
    assert np.allclose(d, want, rtol=1e-14)
    sv = np.linalg.svd(mat(a, n, n), compute_uv=False)
    assert np.allclose(np.sort(sv)[::-1], want, rtol=1e-12)


# ------------------------------------------------------------------------------------------
# dgemm: fixtures, the exact integer case, and the definition on every blocking branch
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["dgemm_tn", "dgemm_nt", "dgemm_tt"])
@pytest.mark.parametrize("avx", [1, 0])
def test_dgemm_variants_match_cblas_fixtures(name, avx):
    f = FIX[name]
    O.set_avx(avx)
    try:
        fn = getattr(O, name)
        for (alpha, beta), want in zip(f["alpha_beta"], (f["want"], f["want_ab"])):
            got = fn(alpha, f["a"], f["b"], beta, f["c0"], f["m"], f["n"], f["k"])
            scale = np.maximum(np.abs(want), 1.0)
            assert np.all(np.abs(got - np.array(want)) / scale < 1e-13)
    finally:
        O.set_avx(-1)


def test_dgemm_tn_exact_on_integers():
    f = FIX["dgemm_tn_exact"]
    got = O.dgemm_tn(1.0, f["a"], f["b"], 0.0, np.zeros(12), f["m"], f["n"], f["k"])
    assert np.array_equal(got, np.array(f["want"]))


def test_dgemm_hand_computation_and_beta():
    # src/blas.rs:828-849
    a = np.array([1.0, 4.0, 2.0, 5.0, 3.0, 6.0])
    b = np.array([7.0, 9.0, 11.0, 8.0, 10.0, 12.0])
    assert list(O.dgemm(1.0, a, b, 0.0, np.zeros(4), 2, 2, 3)) == [58.0, 139.0, 64.0, 154.0]
    assert list(O.dgemm(1.0, a, b, 2.0, np.ones(4), 2, 2, 3)) == [60.0, 141.0, 66.0, 156.0]


@pytest.mark.parametrize("m,n,k", SWEEP_SHAPES)
@pytest.mark.parametrize("avx", [1, 0])
def test_gemm_all_variants_match_the_definition(m, n, k, avx):
    # src/blas.rs:747-794 / :2264-2319: the reference transposes explicitly and multiplies.
    a_op, b_op, c0 = sweep_operands(m, n, k)
    A, B, C0 = mat(a_op, m, k), mat(b_op, k, n), mat(c0, m, n)
    a_t, b_t = transpose_flat(a_op, m, k), transpose_flat(b_op, k, n)
    O.set_avx(avx)
    try:
        for ta in (False, True):
            for tb in (False, True):
                a_st = a_t if ta else a_op
                b_st = b_t if tb else b_op
                for alpha, beta in SWEEP_ALPHA_BETA:
                    want = alpha * (A @ B) + (0.0 if beta == 0.0 else beta * C0)
                    got = mat(O.gemm(ta, tb, alpha, a_st, b_st, beta, c0, m, n, k), m, n)
                    assert np.all(np.abs(got - want) / np.maximum(np.abs(want), 1.0) < 1e-11), (ta, tb, alpha, beta)
    finally:
        O.set_avx(-1)


def test_gemm_beta_zero_ignores_nan_in_c_and_alpha_zero_skips_products():
    a = np.array([np.inf, 1.0, 1.0, 1.0])
    c = np.array([np.nan] * 4)
    got = O.dgemm(1.0, np.ones(4), np.ones(4), 0.0, c, 2, 2, 2)
    assert np.array_equal(got, np.full(4, 2.0))
    got = O.dgemm(0.0, a, np.ones(4), 2.0, np.ones(4), 2, 2, 2)  # inf never multiplied
    assert np.array_equal(got, np.full(4, 2.0))


# ------------------------------------------------------------------------------------------
# dsyrk, triangular solves
# ------------------------------------------------------------------------------------------

def test_dsyrk_both_triangles_and_the_other_is_untouched():
    # src/blas.rs:851-872
    a = np.array([1.0, 3.0, 2.0, 4.0])
    lo = O.dsyrk(1.0, a, 0.0, np.zeros(4), 2, 2, False)
    assert list(lo) == [5.0, 11.0, 0.0, 25.0]
    up = O.dsyrk(1.0, a, 0.0, np.zeros(4), 2, 2, True)
    assert list(up) == [5.0, 0.0, 11.0, 25.0]


@pytest.mark.parametrize("upper", [False, True])
def test_dsyrk_and_rank_match_the_definition(upper):
    n, k = 7, 5
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, n * k)
    c0 = rng.uniform(-1, 1, n * n)
    A, C0 = mat(a, n, k), mat(c0, n, n)
    want = 2.5 * A @ A.T - 1.5 * C0
    got = mat(O.dsyrk(2.5, a, -1.5, c0, n, k, upper), n, n)
    tri = np.triu(np.ones((n, n), bool)) if upper else np.tril(np.ones((n, n), bool))
    assert np.allclose(got[tri], want[tri], rtol=1e-13, atol=1e-13)
    assert np.array_equal(got[~tri], C0[~tri])
    # dsyrk_rank(transposed): A is k x n, C = alpha A^T A + beta C (src/blas.rs:1224-1251)
    at = transpose_flat(a, n, k)
    got2 = mat(O.dsyrk_rank(2.5, at, -1.5, c0, n, k, upper, True), n, n)
    assert np.allclose(got2[tri], want[tri], rtol=1e-13, atol=1e-13)


def test_dtrsv_and_dtrsm_solve_known_systems():
    # src/blas.rs:2761-2821
    x = np.array([1.0, -2.0, 0.5])
    for a, lower in ((lower3(), True), (upper3(), False)):
        b = mat(a, 3, 3) @ x
        assert np.allclose(O.dtrsv(a, b, 3, lower, False), x, atol=1e-14)
    unit = np.array([7.0, 1.0, 4.0, 0.0, 7.0, 5.0, 0.0, 0.0, 7.0])
    b = np.array([[1, 0, 0], [1, 1, 0], [4, 5, 1.0]]) @ x
    assert np.allclose(O.dtrsv(unit, b, 3, True, True), x, atol=1e-14)
    b = mat(lower3(), 3, 3).T @ x
    assert np.allclose(O.dtrsv_transpose(lower3(), b, 3, True, False), x, atol=1e-14)
    X = np.array([1.0, 2.0, 3.0, -1.0, 0.5, 2.0])
    B = cm(mat(lower3(), 3, 3) @ mat(X, 3, 2))
    assert np.allclose(O.dtrsm(2.0, lower3(), B, 3, 2, True, True, False), 2.0 * X, atol=1e-14)
    X = np.array([1.0, 4.0, 2.0, 5.0, 3.0, 6.0])
    B = cm(mat(X, 2, 3) @ mat(lower3(), 3, 3))
    assert np.allclose(O.dtrsm(2.0, lower3(), B, 2, 3, False, True, False), 2.0 * X, atol=1e-14)


@pytest.mark.parametrize("left", [True, False])
@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("unit", [True, False])
def test_dtrsm_matches_scipy(left, lower, unit):
    m, n = 9, 6
    rng = np.random.default_rng(11)
    na = m if left else n
    T = rng.uniform(-1, 1, (na, na)) + 4 * np.eye(na)
    B = rng.uniform(-1, 1, (m, n))
    want = sla.blas.dtrsm(1.5, T, B, side=0 if left else 1, lower=int(lower), diag=int(unit))
    got = mat(O.dtrsm(1.5, cm(T), cm(B), m, n, left, lower, unit), m, n)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------------------------------------
# LU / solve / Cholesky drivers
# ------------------------------------------------------------------------------------------

def test_dgetrf_determinant_by_cofactors():
    # src/lapack/mod.rs:483-495
    st, lu, ipiv, det = O.dgetrf(sample4(), 4)
    assert st == 0 and len(ipiv) == 4
    # the cofactor expansion over the first row of the reference test evaluates to 209
    expect = (4.0 * (3.0 * (5.0 * 6.0 - 2.0 * 2.0) + 1.0 * (0.0 * 2.0 - 5.0 * 1.0))
              - 1.0 * (1.0 * (5.0 * 6.0 - 2.0 * 2.0) + 1.0 * (2.0 * 2.0 - 5.0 * 1.0))
              + 2.0 * (1.0 * (0.0 * 6.0 - 2.0 * 1.0) - 3.0 * (2.0 * 6.0 - 2.0 * 1.0) + 1.0 * (2.0 * 1.0))
              - 1.0 * (1.0 * (0.0 * 2.0 - 5.0 * 1.0) - 3.0 * (2.0 * 2.0 - 5.0 * 1.0)))
    assert expect == 209.0 and abs(det - expect) < 1e-12


def test_dgetrf_singular_and_threshold():
    # src/lapack/mod.rs:497-502 (needs unfused arithmetic), :670-683
    assert O.dgetrf(singular3(), 3)[0] == 1
    st, _, _, det = O.dgetrf(np.array([1e-16, 0.0, 0.0, 1.0]), 2)
    assert st == 0 and abs(det) < 1e-15
    assert O.dgetrf(np.array([1.0, 2.0, 2.0, 4.0]), 2)[0] == 1
    st, _, ipiv, det = O.dgetrf(np.zeros(0), 0)
    assert st == 0 and det == 1.0 and len(ipiv) == 0


def test_dgetrf_rect_reconstructs_a_tall_panel():
    # src/lapack/mod.rs:723-733
    a0 = np.array([1.0, 5.0, 2.0, 3.0, 4.0, 2.0, -1.0, 6.0, 0.0, 1.0, 3.0, 2.0, 1.0, 7.0, -2.0])
    st, a, ipiv = O.dgetrf_rect(a0, 5, 3)
    assert st == 0 and ipiv[0] == 1
    assert O.get01(a0, a, ipiv, 5, 3) < 30
    assert O.dgetrf_rect(np.array([0.0, 0.0, 2.0, 3.0]), 2, 2)[0] == 1
    assert O.dgetrf_rect(np.zeros(4), 1, 2)[0] == 2


def test_dgesv_recovers_a_known_solution_and_reports_singular():
    # src/lapack/mod.rs:537-558
    x = np.array([1.0, -2.0, 3.0, 5.0])
    b = mat(sample4(), 4, 4) @ x
    st, _, got = O.dgesv(sample4(), b)
    assert st == 0 and np.allclose(got, x, atol=1e-12)
    assert O.dgesv(singular3(), np.ones(3))[0] == 1


def test_dpotrf_known_factor_and_not_spd():
    # src/lapack/mod.rs:560-607
    st, a = O.dpotrf(np.array([4.0, 2.0, 2.0, 3.0]), 2)
    assert st == 0 and abs(a[0] - 2) < 1e-14 and abs(a[1] - 1) < 1e-14 and abs(a[3] - math.sqrt(2)) < 1e-14
    assert a[2] == 2.0, "the upper triangle is untouched"
    assert O.dpotrf(np.array([1.0, 2.0, 2.0, 1.0]), 2)[0] == 1
    assert O.dpotrf(np.array([1.0, 1.0, 1.0, 1.0]), 2)[0] == 1


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 9, 33])
def test_drivers_agree_with_scipy_lapack_on_parity_samples(n):
    # src/parity.rs:152-283 at TOL = 1e-10 (src/parity.rs:29-35); scipy stands in for LAPACK.
    a = sample(n)
    st, lu, ipiv, det = O.dgetrf(a, n)
    lu_s, piv_s = sla.lu_factor(mat(a, n, n))
    assert st == 0
    assert np.array_equal(ipiv, piv_s)
    assert np.allclose(mat(lu, n, n), lu_s, rtol=1e-10, atol=1e-10)
    assert abs(det - np.linalg.det(mat(a, n, n))) <= 1e-10 * abs(det)
    x_true = 1.0 + 0.25 * np.arange(n)
    b = mat(a, n, n) @ x_true
    st, _, x = O.dgesv(a, b)
    assert np.allclose(x, sla.solve(mat(a, n, n), b), rtol=1e-10) and np.allclose(x, x_true, rtol=1e-10)
    s = spd(n)
    st, l = O.dpotrf(s, n)
    L = np.tril(mat(l, n, n))
    assert st == 0 and np.allclose(L, sla.cholesky(mat(s, n, n), lower=True), rtol=1e-10, atol=1e-12)


def test_lin_checkers_on_correct_solves_and_hilbert():
    # src/lapack/checkers.rs:307-352
    for n in (1, 2, 5, 8):
        a0 = spd(n)
        x = np.ones(n)
        st, lu, ipiv, _ = O.dgetrf(a0, n)
        assert O.get01(a0, lu, ipiv, n, n) <= 20
        st, ainv = O.dgetri(lu, n, ipiv)
        r3, rcond = O.get03(a0, ainv, n)
        st, _, xs = O.dgesv(a0, mat(a0, n, n) @ x)
        assert r3 <= 20 and O.get04(xs, x, n, 1, rcond) <= 20
    n = 6
    h = hilbert(n)
    st, _, xs = O.dgesv(h, mat(h, n, n) @ np.ones(n))
    st, lu, ipiv, _ = O.dgetrf(h, n)
    _, ainv = O.dgetri(lu, n, ipiv)
    _, rcond = O.get03(h, ainv, n)
    assert rcond < 1e-6 and O.get04(xs, np.ones(n), n, 1, rcond) <= 20
    assert O.get04(xs, np.ones(n), n, 1, 1.0) > 1.0
    for n in (1, 2, 5, 9):
        s = spd(n)
        st, lu, ipiv, _ = O.dgetrf(s, n)
        _, ainv = O.dgetri(lu, n, ipiv)
        r, rc = O.pot03(s, ainv, n)
        assert rc > 0 and r <= 20


def test_checkers_flag_planted_mistakes():
    """A test that cannot fail pins nothing: each planted error must push its ratio far above 30."""
    n = 40
    a = O.larnv(O.SYMMETRIC, SEEDY, n * n)[0]
    st, lu, ipiv, _ = O.dgetrf(a, n)
    assert O.get01(a, lu, ipiv, n, n) < 30
    bad = lu.copy()
    bad[5 + 2 * n] = -bad[5 + 2 * n]  # a multiplier with the wrong sign
    assert O.get01(a, bad, ipiv, n, n) > 1e6
    swapped = ipiv.copy()
    swapped[[3, 7]] = swapped[[7, 3]]  # pivots applied in the wrong order
    assert O.get01(a, lu, swapped, n, n) > 1e6
    x = np.ones(n)
    b = mat(a, n, n) @ x
    assert O.backward_error(a, x, b, n) < 1e-15
    assert O.backward_error(a, x + 1e-9, b, n) > 1e-11


def test_backward_error_is_compensated():
    # b = A x with cancellation: the compensated residual of the exact solution is 0.
    a = cm([[1e16, -1e16], [1.0, 1.0]])
    x = np.array([1.0, 1.0])
    b = np.array([0.0, 2.0])
    assert O.backward_error(a, x, b, 2) == 0.0


# ------------------------------------------------------------------------------------------
# level 1/2 and norms: idamax, dnrm2, dgemv, dgemv_trans, dlange (src/blas.rs:49-158, :628-657)
# ------------------------------------------------------------------------------------------

def _a23():
    """src/blas.rs:669-671: [[1,2,3],[4,5,6]] column-major."""
    return np.array([1.0, 4.0, 2.0, 5.0, 3.0, 6.0])


def _real_matrix(m, n):
    """src/blas.rs:2377-2379."""
    return np.array([float((i * 7) % 19) - 9.0 + i * 0.125 for i in range(m * n)])


# Plausible mistakes in an idamax restatement; each pin vector below must tell at least one of
# them from the crate's rule (strict '>' from -inf: first maximum, NaN never wins).
def _idamax_last_max(x):
    best, ba = 0, -math.inf
    for i, v in enumerate(x):
        if abs(v) >= ba:
            best, ba = i, abs(v)
    return best


def _idamax_nan_wins(x):
    best, ba = 0, -math.inf
    for i, v in enumerate(x):
        if not abs(v) <= ba:
            best, ba = i, abs(v)
    return best


def _idamax_start_at_first(x):
    # seeds the scan with x[0] instead of -inf: a NaN at index 0 then blocks every later entry
    if len(x) == 0:
        return 0
    best, ba = 0, abs(x[0])
    for i, v in enumerate(x):
        if abs(v) > ba:
            best, ba = i, abs(v)
    return best


IDAMAX_PINS = [
    ([], 0),                                  # empty slice yields 0 (src/blas.rs:75-76)
    ([1.0, -3.0, 3.0, 2.0], 1),               # ties: the first maximum wins
    ([-2.0, 5.0, -5.0, 5.0], 1),
    ([1.0, math.nan, 0.5], 0),                # a later NaN is skipped
    ([math.nan, 2.0, -7.0], 2),               # a NaN at index 0 does not block later entries
    ([math.nan, math.nan], 0),                # nothing beats -inf: index 0
    ([0.0, -0.0, 0.0], 0),                    # signed zeros tie
    ([3.0, -math.inf, math.inf], 1),          # infinities compare by magnitude, first wins
]


@pytest.mark.parametrize("x,want", IDAMAX_PINS)
def test_idamax_first_maximum_and_nan_rule(x, want):
    assert O.idamax(np.array(x, dtype=np.float64)) == want


def test_idamax_pins_reject_plausible_mistakes():
    for mutant in (_idamax_last_max, _idamax_nan_wins, _idamax_start_at_first):
        assert any(mutant(x) != want for x, want in IDAMAX_PINS), mutant.__name__
    rng = np.random.default_rng(3)
    for n in (1, 2, 17, 1000):
        x = rng.integers(-4, 5, size=n).astype(np.float64)  # many ties
        assert O.idamax(x) == int(np.argmax(np.abs(x)))     # numpy's argmax is first-max too


def test_dnrm2_closed_forms_and_scaling():
    # src/blas.rs:691-692 and :697-708
    assert O.dnrm2(np.array([3.0, 4.0])) == 5.0
    assert O.dnrm2(np.array([])) == 0.0
    big = O.dnrm2(np.array([1e200, 1e200]))
    assert math.isfinite(big) and abs(big / (1e200 * math.sqrt(2.0)) - 1.0) < 1e-12
    tiny = O.dnrm2(np.array([1e-200, 1e-200]))
    assert tiny > 0.0 and abs(tiny / (1e-200 * math.sqrt(2.0)) - 1.0) < 1e-12
    # a naive sum of squares fails both of the last two; the scaled one does not
    with np.errstate(over="ignore", under="ignore"):
        naive = [np.sqrt(np.sum(np.array([v, v]) * np.array([v, v]))) for v in (1e200, 1e-200)]
    assert not np.isfinite(naive[0]) and naive[1] == 0.0
    assert math.isnan(O.dnrm2(np.array([1.0, math.nan])))     # a NaN propagates
    assert O.dnrm2(np.array([1.0, math.inf])) == math.inf
    rng = np.random.default_rng(5)
    for n in (1, 7, 1000):
        x = rng.standard_normal(n)
        assert abs(O.dnrm2(x) - np.linalg.norm(x)) <= 1e-14 * np.linalg.norm(x)


def test_dgemv_hand_values_and_definition():
    # src/blas.rs:711-723 on the 2x3 matrix a23
    a = _a23()
    assert list(O.dgemv(1.0, a, [1.0, 1.0, 1.0], 0.0, np.zeros(2), 2, 3)) == [6.0, 15.0]
    assert list(O.dgemv(2.0, a, [1.0, 1.0, 1.0], 3.0, np.ones(2), 2, 3)) == [2 * 6.0 + 3, 2 * 15.0 + 3]
    # a distinct x catches a transposed operand or a wrong column stride
    assert list(O.dgemv(1.0, a, [1.0, 10.0, 100.0], 0.0, np.zeros(2), 2, 3)) == [321.0, 654.0]
    # beta = 0 overwrites y (a NaN there does not survive); x_j = 0 skips column j
    y = O.dgemv(1.0, np.array([1.0, 2.0, math.nan, math.nan]), [1.0, 0.0], 0.0, np.array([math.nan] * 2), 2, 2)
    assert list(y) == [1.0, 2.0]
    rng = np.random.default_rng(7)
    for m, n in ((1, 1), (5, 3), (3, 5), (64, 33)):
        A, x, y0 = rng.standard_normal((m, n)), rng.standard_normal(n), rng.standard_normal(m)
        got = O.dgemv(-1.25, cm(A), x, 0.5, y0, m, n)
        want = -1.25 * (A @ x) + 0.5 * y0
        assert np.allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(A).sum())


def test_dgemv_trans_matches_the_definition():
    # src/blas.rs:2410-2432: a 2x3, x 2 entries, y 3 entries
    m, n = 2, 3
    a, x = _real_matrix(m, n), _real_matrix(m, 1)
    for alpha, beta in ((1.0, 0.0), (-2.5, 1.75)):
        y0 = _real_matrix(1, n)
        want = np.array([alpha * sum(a[i + j * m] * x[i] for i in range(m)) + beta * y0[j] for j in range(n)])
        got = O.dgemv_trans(alpha, a, x, beta, y0, m, n)
        assert np.all(np.abs(got - want) / np.maximum(np.abs(want), 1.0) < 1e-13)
    # beta = 0 ignores a NaN in y; the transposed product differs from the plain one here
    got = O.dgemv_trans(1.0, _a23(), [1.0, 10.0], 0.0, np.array([math.nan] * 3), 2, 3)
    assert list(got) == [41.0, 52.0, 63.0]
    rng = np.random.default_rng(8)
    for m, n in ((1, 4), (6, 2), (70, 31)):
        A, x = rng.standard_normal((m, n)), rng.standard_normal(m)
        assert np.allclose(O.dgemv_trans(1.0, cm(A), x, 0.0, np.zeros(n), m, n), A.T @ x, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("norm,want", [("M", 6.0), ("m", 6.0), ("1", 9.0), ("O", 9.0), ("o", 9.0), ("I", 15.0),
                                       ("i", 15.0), ("F", math.sqrt(91.0)), ("f", math.sqrt(91.0)),
                                       ("E", math.sqrt(91.0)), ("e", math.sqrt(91.0))])
def test_dlange_every_norm_on_the_crate_matrix(norm, want):
    # src/blas.rs:875-887; the one- and infinity-norms differ on a23, so swapped loops fail
    got = O.dlange(norm, _a23(), 2, 3)
    assert abs(got - want) <= 1e-12 * want


def test_dlange_against_numpy_and_nan_rule():
    rng = np.random.default_rng(9)
    A = rng.standard_normal((7, 4))
    a = cm(A)
    assert O.dlange("M", a, 7, 4) == np.abs(A).max()
    assert abs(O.dlange("1", a, 7, 4) - np.linalg.norm(A, 1)) <= 1e-14 * np.linalg.norm(A, 1)
    assert abs(O.dlange("I", a, 7, 4) - np.linalg.norm(A, np.inf)) <= 1e-14 * np.linalg.norm(A, np.inf)
    assert abs(O.dlange("F", a, 7, 4) - np.linalg.norm(A, "fro")) <= 1e-14 * np.linalg.norm(A, "fro")
    # f64::max ignores a NaN operand, so 'M' skips NaN entries; 'F' goes through dnrm2 (NaN)
    b = np.array([math.nan, -3.0, 2.0, 1.0])
    assert O.dlange("M", b, 2, 2) == 3.0
    assert math.isnan(O.dlange("F", b, 2, 2))
    assert O.dlange("M", np.zeros(0), 0, 0) == 0.0
    assert math.isnan(O.dlange("X", b, 2, 2))   # the crate panics; the oracle returns NaN


def test_laruv_laran_are_the_uniform_stream():
    # src/lapack/families.rs:230-260: DLARNV's uniform case copies DLARUV, and DLARAN is DLARUV
    # with n = 1; both advance the seed by the draws taken
    x, s1 = O.laruv(SEEDY, 40)
    y, s2 = O.larnv(O.UNIFORM, SEEDY, 40)
    assert np.array_equal(x, y) and list(s1) == list(s2)
    seed = np.array(SEEDY, dtype=np.uint32)
    for i in range(5):
        v = O.lib().or_laran(seed.ctypes.data_as(O._c_u32_p))
        assert v == x[i]
    assert np.all((x > 0.0) & (x < 1.0))


@pytest.mark.parametrize("mode", [1, 2, 3, 4, -3])
def test_latm1_closed_form_spectra(mode):
    # src/lapack/families.rs:262-356: modes 1-4 are closed forms of cond; a negative mode reverses
    n, cond = 9, 1e4
    info, d, _ = O.latm1(mode, cond, 0, O.UNIFORM, SEEDY, np.zeros(n))
    assert info == 0
    i = np.arange(n)
    want = {1: np.where(i == 0, 1.0, 1.0 / cond), 2: np.where(i == n - 1, 1.0 / cond, 1.0),
            3: cond ** (-i / (n - 1.0)), 4: 1.0 - i / (n - 1.0) * (1.0 - 1.0 / cond)}[abs(mode)]
    if mode < 0:
        want = want[::-1]
    assert np.allclose(d, want, rtol=1e-14, atol=1e-15)
    assert O.latm1(7, cond, 0, O.UNIFORM, SEEDY, np.zeros(n))[0] == -1
    assert O.latm1(3, 0.5, 0, O.UNIFORM, SEEDY, np.zeros(n))[0] == -3


def test_dgetrs_applies_the_pivots_and_solves():
    # src/lapack/mod.rs:537-558: x = [1, -2, 3, 5] through dgetrf + dgetrs, one call per column
    a = sample4()
    x = np.array([1.0, -2.0, 3.0, 5.0])
    b = mat(a, 4, 4) @ x
    st, lu, ipiv, _ = O.dgetrf(a, 4)
    assert st == 0
    assert np.allclose(O.dgetrs(lu, ipiv, b), x, atol=1e-14)
    # pivots that really swap: a permuted identity's LU is trivial, dgetrs must undo P
    p = np.array([2, 0, 3, 1])
    A = np.eye(4)[p] * np.array([2.0, 3.0, 5.0, 7.0])[:, None]
    st, lu, ipiv, _ = O.dgetrf(cm(A), 4)
    assert st == 0 and list(ipiv) != [0, 1, 2, 3]
    b = A @ x
    assert np.allclose(O.dgetrs(lu, ipiv, b), x, atol=1e-14)


# ------------------------------------------------------------------------------------------
# dsytrd's dsyr2k (src/lapack/sym.rs:582-668)
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("upper", [True, False])
@pytest.mark.parametrize("m,k", [(1, 1), (5, 3), (17, 9), (40, 64)])
def test_dsyr2k_matches_the_definition_on_one_triangle(upper, m, k):
    rng = np.random.default_rng(m * 100 + k)
    V, W, C = rng.standard_normal((m, k)), rng.standard_normal((m, k)), rng.standard_normal((m, m))
    tri = np.triu(np.ones((m, m), bool)) if upper else np.tril(np.ones((m, m), bool))
    for alpha, beta in ((-1.0, 1.0), (0.75, -2.0), (1.5, 0.0)):
        got = mat(O.dsyr2k(upper, m, k, alpha, cm(V), cm(W), beta, cm(C)), m, m)
        want = alpha * (V @ W.T) + alpha * (W @ V.T) + beta * C
        bound = 4 * k * np.finfo(float).eps * (abs(alpha) * 2 * np.abs(V) @ np.abs(W).T + abs(beta) * np.abs(C))
        assert np.all(np.abs(got - want)[tri] <= bound[tri] + 1e-300)
        assert np.array_equal(got[~tri], C[~tri])   # the other triangle is untouched
    # a plausible mistake, 2 alpha V W^T (one product counted twice), breaks the definition
    wrong = 2 * (V @ W.T)
    assert m == 1 or not np.allclose(wrong[tri], (V @ W.T + W @ V.T)[tri])


def test_dsyr2k_beta_zero_and_the_empty_cases():
    m, k = 4, 2
    rng = np.random.default_rng(1)
    V, W = rng.standard_normal((m, k)), rng.standard_normal((m, k))
    C = np.full((m, m), np.nan)
    got = mat(O.dsyr2k(False, m, k, 1.0, cm(V), cm(W), 0.0, cm(C)), m, m)
    tri = np.tril(np.ones((m, m), bool))
    assert np.all(np.isfinite(got[tri])) and np.all(np.isnan(got[~tri]))   # beta = 0 does not read C
    C = rng.standard_normal((m, m))
    # k = 0: only alpha = 0 touches C (beta scales the triangle, beta = 0 clears it)
    assert np.array_equal(O.dsyr2k(True, m, 0, 1.0, np.zeros(0), np.zeros(0), 3.0, cm(C)), cm(C))
    got = mat(O.dsyr2k(True, m, 0, 0.0, np.zeros(0), np.zeros(0), 3.0, cm(C)), m, m)
    up = np.triu(np.ones((m, m), bool))
    assert np.array_equal(got[up], 3.0 * C[up]) and np.array_equal(got[~up], C[~up])
    got = mat(O.dsyr2k(False, m, 0, 0.0, np.zeros(0), np.zeros(0), 0.0, cm(C)), m, m)
    assert np.all(got[tri] == 0.0) and np.array_equal(got[~tri], C[~tri])
