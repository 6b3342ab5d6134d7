"""GPU parity of the level-3 routines (dgemm x4, dsyrk, dtrsm) against the CPU oracle.

Bars (each stated where it is asserted):
  * the crate's captured CBLAS fixtures: 1e-13 relative (src/blas.rs:1669-1676);
  * the small-integer dgemm_tn case: bit-exact (src/blas.rs:2008-2022);
  * everything else: elementwise |C - C_oracle| <= 4*k*eps*(|alpha|(|A||B|)_ij + |beta||C0_ij|)
    (twice the accumulated-rounding bound of either side), plus the crate's own 1e-11
    relative sweep bar against the definition.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from matrices import (SWEEP_ALPHA_BETA, SWEEP_SHAPES, cm, gemm_error_bound, lower3, mat, sweep_operands,
                      transpose_flat)

pytestmark = pytest.mark.gpu

al = pytest.importorskip("algebra_linear_amd")
blas = al.blas

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
SEEDY = [1988, 1989, 1990, 1991]
GEMMS = {(False, False): "dgemm", (True, False): "dgemm_tn", (False, True): "dgemm_nt", (True, True): "dgemm_tt"}


def lapack_uniform(count, seed=SEEDY):
    return O.larnv(O.SYMMETRIC, seed, count)[0]


@pytest.mark.parametrize("name", ["dgemm_tn", "dgemm_nt", "dgemm_tt"])
def test_gemm_variants_match_cblas_fixtures(name):
    f = FIX[name]
    for (alpha, beta), want in zip(f["alpha_beta"], (f["want"], f["want_ab"])):
        c = np.array(f["c0"])
        getattr(blas, name)(alpha, f["a"], f["b"], beta, c, f["m"], f["n"], f["k"])
        scale = np.maximum(np.abs(want), 1.0)
        assert np.all(np.abs(c - np.array(want)) / scale < 1e-13), name


def test_gemm_tn_exact_on_integers():
    f = FIX["dgemm_tn_exact"]
    c = np.zeros(12)
    blas.dgemm_tn(1.0, f["a"], f["b"], 0.0, c, f["m"], f["n"], f["k"])
    assert np.array_equal(c, np.array(f["want"]))


def test_gemm_hand_computation_and_beta():
    a = np.array([1.0, 4.0, 2.0, 5.0, 3.0, 6.0])
    b = np.array([7.0, 9.0, 11.0, 8.0, 10.0, 12.0])
    c = np.zeros(4)
    blas.dgemm(1.0, a, b, 0.0, c, 2, 2, 3)
    assert list(c) == [58.0, 139.0, 64.0, 154.0]
    c = np.ones(4)
    blas.dgemm(1.0, a, b, 2.0, c, 2, 2, 3)
    assert list(c) == [60.0, 141.0, 66.0, 156.0]


def test_gemm_shape_panics_with_the_crates_messages():
    with pytest.raises(al.BlasPanic, match="dgemm: a must be m\\*k"):
        blas.dgemm(1.0, np.zeros(5), np.zeros(6), 0.0, np.zeros(4), 2, 2, 3)
    with pytest.raises(al.BlasPanic, match="dgemm_tn: b must be k\\*n"):
        blas.dgemm_tn(1.0, np.zeros(6), np.zeros(5), 0.0, np.zeros(4), 2, 2, 3)
    with pytest.raises(al.BlasPanic, match="dgemm_nt: c must be m\\*n"):
        blas.dgemm_nt(1.0, np.zeros(6), np.zeros(6), 0.0, np.zeros(5), 2, 2, 3)


def _check_gemm(ta, tb, m, n, k, alpha, beta, a_op, b_op, c0):
    A, B, C0 = mat(a_op, m, k), mat(b_op, k, n), mat(c0, m, n)
    a_st = transpose_flat(a_op, m, k) if ta else a_op
    b_st = transpose_flat(b_op, k, n) if tb else b_op
    c = c0.copy()
    getattr(blas, GEMMS[(ta, tb)])(alpha, a_st, b_st, beta, c, m, n, k)
    want = O.gemm(ta, tb, alpha, a_st, b_st, beta, c0, m, n, k)
    bound = gemm_error_bound(alpha, A, B, beta, C0, k)
    err = np.abs(mat(c, m, n) - mat(want, m, n))
    assert np.all(err <= bound), (ta, tb, m, n, k, alpha, beta, float((err / bound).max()))
    return mat(c, m, n), A, B, C0


@pytest.mark.parametrize("m,n,k", SWEEP_SHAPES)
def test_gemm_sweep_all_variants(m, n, k):
    """The crate's branch-reaching shapes (src/blas.rs:747-794, :2264-2319), every variant."""
    a_op, b_op, c0 = sweep_operands(m, n, k)
    for (ta, tb) in GEMMS:
        for alpha, beta in SWEEP_ALPHA_BETA:
            got, A, B, C0 = _check_gemm(ta, tb, m, n, k, alpha, beta, a_op, b_op, c0)
            want = alpha * (A @ B) + (0.0 if beta == 0.0 else beta * C0)
            assert np.all(np.abs(got - want) / np.maximum(np.abs(want), 1.0) < 1e-11)


@pytest.mark.parametrize("m,n,k", [(512, 512, 512), (1000, 777, 333), (129, 1031, 257), (1024, 1024, 1024)])
def test_gemm_uniform_inputs_all_variants(m, n, k):
    """Config-1/2 style inputs: uniform(-1,1) from the LAPACK stream, A then B then C."""
    ab = lapack_uniform(m * k + k * n + m * n)
    a_op, b_op, c0 = ab[: m * k], ab[m * k: m * k + k * n], ab[m * k + k * n:]
    for (ta, tb) in GEMMS:
        for alpha, beta in ((1.0, 0.0), (2.5, -1.5)):
            _check_gemm(ta, tb, m, n, k, alpha, beta, a_op, b_op, c0)


def test_gemm_beta_zero_overwrites_nan_and_alpha_zero_never_multiplies():
    c = np.full(4, np.nan)
    blas.dgemm(1.0, np.ones(4), np.ones(4), 0.0, c, 2, 2, 2)
    assert np.array_equal(c, np.full(4, 2.0))
    c = np.ones(4)
    blas.dgemm(0.0, np.array([np.inf, 1, 1, 1.0]), np.ones(4), 2.0, c, 2, 2, 2)
    assert np.array_equal(c, np.full(4, 2.0))
    c = np.ones(6)
    blas.dgemm(1.0, np.zeros(0), np.zeros(0), 3.0, c, 2, 3, 0)  # k == 0: C := beta*C
    assert np.array_equal(c, np.full(6, 3.0))


# ------------------------------------------------------------------------------------------
# dsyrk
# ------------------------------------------------------------------------------------------

def test_dsyrk_hand_case_leaves_the_other_triangle_alone():
    a = np.array([1.0, 3.0, 2.0, 4.0])
    lo = np.zeros(4)
    blas.dsyrk(1.0, a, 0.0, lo, 2, 2, False)
    assert list(lo) == [5.0, 11.0, 0.0, 25.0]
    up = np.zeros(4)
    blas.dsyrk(1.0, a, 0.0, up, 2, 2, True)
    assert list(up) == [5.0, 0.0, 11.0, 25.0]


@pytest.mark.parametrize("n,k", [(7, 5), (130, 64), (300, 257), (1024, 512)])
@pytest.mark.parametrize("upper", [False, True])
def test_dsyrk_matches_oracle_and_never_reads_the_other_triangle(n, k, upper):
    a = lapack_uniform(n * k)
    c0 = lapack_uniform(n * n, [11, 12, 13, 15])
    C0 = mat(c0, n, n)
    tri = np.triu(np.ones((n, n), bool)) if upper else np.tril(np.ones((n, n), bool))
    poisoned = C0.copy()
    poisoned[~tri] = np.nan  # must be neither read (beta != 0) nor written
    c = cm(poisoned)
    blas.dsyrk(2.5, a, -1.5, c, n, k, upper)
    got = mat(c, n, n)
    want = mat(O.dsyrk(2.5, a, -1.5, c0, n, k, upper), n, n)
    A = mat(a, n, k)
    bound = gemm_error_bound(2.5, A, A.T, -1.5, C0, k)
    assert np.all(np.abs(got[tri] - want[tri]) <= bound[tri])
    assert np.all(np.isnan(got[~tri]))


# ------------------------------------------------------------------------------------------
# dtrsm
# ------------------------------------------------------------------------------------------

def test_dtrsm_solves_both_sides_of_the_crates_example():
    alpha = 2.0
    x = np.array([1.0, 2.0, 3.0, -1.0, 0.5, 2.0])
    b = cm(mat(lower3(), 3, 3) @ mat(x, 3, 2))
    blas.dtrsm(alpha, lower3(), b, 3, 2, True, True, False)
    assert np.allclose(b, alpha * x, atol=1e-14)
    x = np.array([1.0, 4.0, 2.0, 5.0, 3.0, 6.0])
    b = cm(mat(x, 2, 3) @ mat(lower3(), 3, 3))
    blas.dtrsm(alpha, lower3(), b, 2, 3, False, True, False)
    assert np.allclose(b, alpha * x, atol=1e-14)


@pytest.mark.parametrize("m,n", [(3, 2), (64, 64), (150, 70), (257, 300)])
@pytest.mark.parametrize("left", [True, False])
@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("unit", [True, False])
def test_dtrsm_matches_oracle(m, n, left, lower, unit):
    na = m if left else n
    rng = np.random.default_rng(m * 7 + n)
    T = rng.uniform(-1, 1, (na, na)) / np.sqrt(na)
    np.fill_diagonal(T, 1.0 + rng.uniform(0, 1, na))
    other = np.triu(np.ones((na, na), bool), 1) if lower else np.tril(np.ones((na, na), bool), -1)
    Tp = T.copy()
    Tp[other] = np.nan  # the untaken triangle must not be read
    if unit:
        np.fill_diagonal(Tp, np.nan)  # nor a unit diagonal
    B = rng.uniform(-1, 1, (m, n))
    b = cm(B)
    blas.dtrsm(1.5, cm(Tp), b, m, n, left, lower, unit)
    Tc = np.where(other, 0.0, T)
    if unit:
        np.fill_diagonal(Tc, 1.0)
    want = O.dtrsm(1.5, cm(Tc), cm(B), m, n, left, lower, unit)
    X = mat(want, m, n)
    # residual test, normwise: |op(T) X - alpha B| relative to |T||X|
    got = mat(b, m, n)
    resid = (Tc @ got - 1.5 * B) if left else (got @ Tc - 1.5 * B)
    scale = (np.abs(Tc) @ np.abs(got)) if left else (np.abs(got) @ np.abs(Tc))
    eps = np.finfo(float).eps
    assert np.all(np.abs(resid) <= 4 * na * eps * scale + 1e-300)
    assert np.allclose(got, X, rtol=1e-6, atol=1e-9)  # forward error carries cond(T)


def test_dtrsm_shape_panics_and_empty_is_a_noop():
    with pytest.raises(al.BlasPanic, match="dtrsm: b must be m\\*n"):
        blas.dtrsm(1.0, np.zeros(4), np.zeros(3), 2, 2, True, True, False)
    with pytest.raises(al.BlasPanic, match="dtrsm: a must be m\\*m for the left side"):
        blas.dtrsm(1.0, np.zeros(3), np.zeros(4), 2, 2, True, True, False)
    blas.dtrsm(1.0, np.zeros(0), np.zeros(0), 0, 0, True, True, False)
