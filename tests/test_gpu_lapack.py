"""GPU parity of the LU / solve / Cholesky drivers against the CPU oracle.

Bars:
  * n <= 64 (one 64-wide panel / one Cholesky block): the GPU performs the crate's own
    operations in the crate's order without FMA, so LU factors, pivots, determinant and the
    Cholesky factor are compared BIT FOR BIT with the oracle;
  * larger n: normwise backward error of the factorization ||PA - LU||_1/(n eps ||A||_1)
    (DGET01) < 30 and ||A - LL^T||_1/(n eps ||A||_1) < 30 (the LAPACK testing bar), the
    solve's normwise backward error ||b - Ax||_inf/(||A|| ||x|| + ||b||) <= 1e-14 (computed
    on the CPU in FP64 with compensated sums), and the pivots agree with the oracle;
  * error paths exactly as the crate's tests (src/lapack/mod.rs:497-616, :670-751).
"""
import math

import numpy as np
import pytest

import oracle as O
from matrices import cm, hilbert, mat, sample, sample4, singular3, spd

pytestmark = pytest.mark.gpu

al = pytest.importorskip("algebra_linear_amd")
lapack = al.lapack
SEEDY = [1988, 1989, 1990, 1991]
EPS = np.finfo(float).eps


def uniform(count, seed=SEEDY):
    return O.larnv(O.SYMMETRIC, seed, count)[0]


def spd_gg(n, seed=SEEDY):
    """Config-2 SPD input: A = G G^T + n I with G uniform(-1,1)."""
    G = mat(uniform(n * n, seed), n, n)
    return cm(G @ G.T + n * np.eye(n))


def chol_residual(a0, l, n):
    L = np.tril(mat(l, n, n))
    A = mat(a0, n, n)
    return np.abs(A - L @ L.T).sum(axis=0).max() / (n * EPS * np.abs(A).sum(axis=0).max())


# ------------------------------------------------------------------------------------------
# the crate's own cases
# ------------------------------------------------------------------------------------------

def test_dgetrf_determinant_and_pivots_of_sample4():
    a = sample4()
    ipiv, det = lapack.dgetrf(a, 4)
    st, lu, piv_o, det_o = O.dgetrf(sample4(), 4)
    assert len(ipiv) == 4 and list(ipiv) == list(piv_o)
    assert abs(det - 209.0) < 1e-12 and det == det_o
    assert np.array_equal(a, lu)


def test_dgetrf_detects_singular_bad_shape_and_tiny_pivot():
    with pytest.raises(al.LapackError) as e:
        lapack.dgetrf(singular3(), 3)
    assert e.value.kind == al.SINGULAR
    with pytest.raises(al.LapackError) as e:
        lapack.dgetrf(np.array([1.0, 2.0, 3.0]), 2)
    assert e.value.kind == al.BAD_SHAPE
    a = np.array([1e-16, 0.0, 0.0, 1.0])
    ipiv, det = lapack.dgetrf(a, 2)
    assert len(ipiv) == 2 and abs(det) < 1e-15
    with pytest.raises(al.LapackError) as e:
        lapack.dgetrf(np.array([1.0, 2.0, 2.0, 4.0]), 2)
    assert e.value.kind == al.SINGULAR


def test_empty_matrices_are_handled():
    assert lapack.dgetrf(np.zeros(0), 0) == ([], 1.0)
    lapack.dgesv(np.zeros(0), np.zeros(0))
    lapack.dpotrf(np.zeros(0), 0)
    lapack.dgetrf_rect(np.zeros(0), 3, 0, np.zeros(0, dtype=np.uintp))


def test_dgesv_recovers_a_known_solution_and_reports_singular():
    x = np.array([1.0, -2.0, 3.0, 5.0])
    b = mat(sample4(), 4, 4) @ x
    a = sample4()
    lapack.dgesv(a, b)
    assert np.allclose(b, x, atol=1e-12)
    with pytest.raises(al.LapackError) as e:
        lapack.dgesv(singular3(), np.ones(3))
    assert e.value.kind == al.SINGULAR


def test_dpotrf_known_factor_and_not_spd():
    a = np.array([4.0, 2.0, 2.0, 3.0])
    lapack.dpotrf(a, 2)
    assert abs(a[0] - 2) < 1e-14 and abs(a[1] - 1) < 1e-14 and abs(a[3] - math.sqrt(2)) < 1e-14
    assert a[2] == 2.0, "upper triangle untouched"
    for bad in (np.array([1.0, 2.0, 2.0, 1.0]), np.array([1.0, 1.0, 1.0, 1.0])):
        with pytest.raises(al.LapackError) as e:
            lapack.dpotrf(bad, 2)
        assert e.value.kind == al.SINGULAR


def test_dgetrf_rect_reconstructs_a_tall_panel_and_rejects_bad_shapes():
    a0 = np.array([1.0, 5.0, 2.0, 3.0, 4.0, 2.0, -1.0, 6.0, 0.0, 1.0, 3.0, 2.0, 1.0, 7.0, -2.0])
    a = a0.copy()
    ipiv = np.zeros(3, dtype=np.uintp)
    lapack.dgetrf_rect(a, 5, 3, ipiv)
    assert ipiv[0] == 1
    st, lu_o, piv_o = O.dgetrf_rect(a0, 5, 3)
    assert np.array_equal(a, lu_o) and list(ipiv) == list(piv_o)
    with pytest.raises(al.LapackError) as e:
        lapack.dgetrf_rect(np.zeros(4), 1, 2, np.zeros(2, dtype=np.uintp))
    assert e.value.kind == al.BAD_SHAPE
    with pytest.raises(al.LapackError) as e:
        lapack.dgetrf_rect(np.zeros(4), 2, 2, np.zeros(1, dtype=np.uintp))
    assert e.value.kind == al.BAD_SHAPE
    with pytest.raises(al.LapackError) as e:
        lapack.dgetrf_rect(np.array([0.0, 0.0, 2.0, 3.0]), 2, 2, np.zeros(2, dtype=np.uintp))
    assert e.value.kind == al.SINGULAR


def test_nan_pivot_semantics_follow_the_crates_scan():
    # a NaN on the diagonal blocks every later candidate (pivot stays k); a later NaN is skipped
    a = cm([[np.nan, 1.0], [5.0, 1.0]])
    ipiv, _ = lapack.dgetrf(a, 2)
    st, _, piv_o, _ = O.dgetrf(cm([[np.nan, 1.0], [5.0, 1.0]]), 2)
    assert ipiv == list(piv_o) and ipiv[0] == 0
    a = cm([[1.0, 1.0, 0.0], [np.nan, 2.0, 1.0], [3.0, 0.0, 1.0]])
    ipiv, _ = lapack.dgetrf(a.copy(), 3)
    st, _, piv_o, _ = O.dgetrf(a, 3)
    assert ipiv[0] == 2 and ipiv[0] == piv_o[0]


# ------------------------------------------------------------------------------------------
# bit-exact inside one panel
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 5, 9, 17, 33, 64])
def test_dgetrf_is_bit_exact_within_one_panel(n):
    a0 = uniform(n * n)
    a = a0.copy()
    ipiv, det = lapack.dgetrf(a, n)
    st, lu, piv_o, det_o = O.dgetrf(a0, n)
    assert st == 0
    assert ipiv == list(piv_o)
    assert np.array_equal(a, lu)
    assert det == det_o or (math.isnan(det) and math.isnan(det_o))


@pytest.mark.parametrize("nrows,ncols", [(5, 3), (100, 64), (1000, 40), (3000, 64)])
def test_dgetrf_rect_is_bit_exact_within_one_panel(nrows, ncols):
    # nrows > 256 spreads the panel over several cooperating workgroups (grid barrier path)
    a0 = uniform(nrows * ncols, [7, 8, 9, 11])
    a = a0.copy()
    ipiv = np.zeros(ncols, dtype=np.uintp)
    lapack.dgetrf_rect(a, nrows, ncols, ipiv)
    st, lu, piv_o = O.dgetrf_rect(a0, nrows, ncols)
    assert list(ipiv) == list(piv_o)
    assert np.array_equal(a, lu)


@pytest.mark.parametrize("nrows,ncols", [(1000, 64), (2600, 300), (3000, 3000), (16384, 64)])
def test_pinned_and_unpinned_panels_agree(nrows, ncols):
    # the cooperative panels give the same pivots and factors whether they run on one half of
    # the chip with their exchange words near it, or anywhere (4 to 64 workgroups, a count that
    # is not a multiple of four included)
    from algebra_linear_amd import _lib

    L = _lib.load()
    half = L.al_debug_panel_pin_half(None)
    assert half in (-1, 0, 1)
    print(f"cooperative panels pinned to half {half}", flush=True)
    a0 = uniform(nrows * ncols, [5, 6, 7, 9])
    a1, a2 = a0.copy(), a0.copy()
    p1, p2 = np.zeros(ncols, dtype=np.uintp), np.zeros(ncols, dtype=np.uintp)
    lapack.dgetrf_rect(a1, nrows, ncols, p1)
    L.al_debug_set_panel_pinning(0)
    try:
        lapack.dgetrf_rect(a2, nrows, ncols, p2)
    finally:
        L.al_debug_set_panel_pinning(1)
    assert list(p1) == list(p2)
    assert np.array_equal(a1, a2)


@pytest.mark.parametrize("zero_rows", [200, 600])
def test_dgetrf_rect_negative_zeros_in_a_workgroups_rows(zero_rows):
    # 600 x 64 runs as three cooperating workgroups of 200 rows; column 0 of the first
    # workgroup's rows (or of every row) holds -0.0, which ties with +0.0 and never beats a
    # nonzero entry: pivots and factors as the crate computes them (or its zero-pivot report)
    nrows, ncols = 600, 64
    a0 = uniform(nrows * ncols, [3, 1, 4, 1])
    a0[:zero_rows] = -0.0
    a = a0.copy()
    ipiv = np.zeros(ncols, dtype=np.uintp)
    st, lu, piv_o = O.dgetrf_rect(a0, nrows, ncols)
    if st != 0:
        with pytest.raises(al.LapackError) as e:
            lapack.dgetrf_rect(a, nrows, ncols, ipiv)
        assert e.value.kind == al.SINGULAR
        return
    lapack.dgetrf_rect(a, nrows, ncols, ipiv)
    assert list(ipiv) == list(piv_o)
    assert np.array_equal(a, lu)


@pytest.mark.parametrize("n", [1, 2, 5, 9, 33, 64, 65, 100, 128])
def test_dpotrf_is_bit_exact_within_one_block(n):
    a0 = spd(n) if n < 10 else spd_gg(n)
    a = a0.copy()
    lapack.dpotrf(a, n)
    st, l = O.dpotrf(a0, n)
    assert st == 0 and np.array_equal(a, l)


# ------------------------------------------------------------------------------------------
# blocked sizes: LIN residuals and backward error
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [65, 200, 300, 512, 1000])
def test_dgetrf_blocked_residual_and_pivots(n):
    a0 = uniform(n * n, [n % 4096, 3, 5, 7])
    a = a0.copy()
    ipiv, det = lapack.dgetrf(a, n)
    r = O.get01(a0, a, ipiv, n, n)
    assert r < 30, r
    st, lu, piv_o, det_o = O.dgetrf(a0, n)
    agree = np.mean(np.array(ipiv) == piv_o)
    assert agree > 0.95, agree  # rounding-level near-ties may pick the other row
    if math.isfinite(det_o) and det_o != 0:
        assert abs(det - det_o) <= 1e-8 * abs(det_o)
    else:
        assert (math.isinf(det) and math.isinf(det_o) and np.sign(det) == np.sign(det_o)) or det == det_o


@pytest.mark.parametrize("nrows,ncols", [(600, 130), (2000, 300), (5000, 257)])
def test_dgetrf_rect_blocked_residual(nrows, ncols):
    a0 = uniform(nrows * ncols, [1, 2, 3, 5])
    a = a0.copy()
    ipiv = np.zeros(ncols, dtype=np.uintp)
    lapack.dgetrf_rect(a, nrows, ncols, ipiv)
    assert O.get01(a0, a, ipiv, nrows, ncols) < 30


@pytest.mark.parametrize("n,nrhs", [(512, 4), (1000, 1), (1500, 16)])
def test_dgesv_backward_error(n, nrhs):
    ab = uniform(n * n + n * nrhs, [9, 9, 9, 9])
    a0, b0 = ab[: n * n], ab[n * n:]
    a, b = a0.copy(), b0.copy()
    ipiv = lapack.dgesv_nrhs(a, b, n, nrhs)
    be = O.backward_error(a0, b, b0, n, nrhs)
    assert be <= 1e-14, be
    assert O.get01(a0, a, ipiv, n, n) < 30


def test_dgesv_single_rhs_and_dgetrs_agree_with_oracle():
    n = 300
    a0 = uniform(n * n, [2, 4, 6, 9])
    x = 1.0 + 0.25 * np.arange(n)
    b0 = mat(a0, n, n) @ x
    a, b = a0.copy(), b0.copy()
    lapack.dgesv(a, b)
    assert O.backward_error(a0, b, b0, n) <= 1e-14
    lu = a0.copy()
    ipiv, _ = lapack.dgetrf(lu, n)
    b2 = b0.copy()
    lapack.dgetrs(lu, ipiv, b2)
    assert np.allclose(b2, b, rtol=1e-12, atol=1e-12)
    bo = O.dgetrs(lu, ipiv, b0)
    assert np.allclose(b2, bo, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("n", [1, 7, 8, 9, 1023, 1024, 1025, 3001])
def test_dgetrs_applies_any_pivot_sequence_in_order(n):
    # the crate's dgetrs swaps b[i] and b[ipiv[i]] for i = 0, 1, ... (src/lapack/mod.rs:330-354)
    # whatever the pivots are, including ipiv[i] < i; with LU = I the solve is exact, so x is b
    # with the interchanges applied, bit for bit (lengths with a remainder of 1 and 7 pivots, and
    # across the 1024-pivot chunks in which the device composes them)
    rng = np.random.default_rng(n)
    ipiv = rng.integers(0, n, size=n)
    lu = np.eye(n).reshape(-1, order="F").copy()
    b0 = rng.standard_normal(n)
    want = b0.copy()
    for i, p in enumerate(ipiv):
        want[i], want[p] = want[p], want[i]
    b = b0.copy()
    lapack.dgetrs(lu, ipiv, b)
    assert np.array_equal(b, want)
    assert np.array_equal(want, O.dgetrs(lu, ipiv, b0))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 9])
def test_dgesv_on_the_parity_samples(n):
    # src/parity.rs:152-195: sample(n), x_true = 1 + i/4, TOL 1e-10
    a = sample(n)
    x = 1.0 + 0.25 * np.arange(n)
    b = mat(a, n, n) @ x
    lapack.dgesv(a.copy(), b)
    assert np.allclose(b, x, rtol=1e-10)


def test_hilbert_forward_error_scaled_by_rcond():
    # src/lapack/checkers.rs:335-352
    n = 6
    h = hilbert(n)
    b = mat(h, n, n) @ np.ones(n)
    lapack.dgesv(h.copy(), b)
    st, lu, ipiv, _ = O.dgetrf(h, n)
    _, ainv = O.dgetri(lu, n, ipiv)
    _, rcond = O.get03(h, ainv, n)
    assert O.get04(b, np.ones(n), n, 1, rcond) <= 20


# sizes around the blocking: 128-wide diagonal blocks inside 384-wide look-ahead blocks
# (129: one column past a diagonal block; 256, 384: whole blocks; 385, 769: one column into the
# next look-ahead block; 512: a partial look-ahead block of whole diagonal blocks)
@pytest.mark.parametrize("n", [65, 129, 200, 256, 300, 384, 385, 512, 700, 769])
def test_dpotrf_blocked_residual_and_upper_untouched(n):
    a0 = spd_gg(n, [n, 1, 1, 1])
    A = mat(a0, n, n)
    P = A.copy()
    P[np.triu(np.ones((n, n), bool), 1)] = np.nan  # never read, never written
    a = cm(P)
    lapack.dpotrf(a, n)
    got = mat(a, n, n)
    assert np.all(np.isnan(got[np.triu(np.ones((n, n), bool), 1)]))
    assert chol_residual(a0, a, n) < 30
    st, l = O.dpotrf(a0, n)
    assert np.allclose(np.tril(got), np.tril(mat(l, n, n)), rtol=1e-10, atol=1e-12)


def test_dpotrf_extreme_pivot_scales_in_later_blocks():
    # a diagonal matrix whose entries span 1e-305 .. 1e305 (past the blocks that use the crate's
    # own arithmetic): L is the entrywise square root, to an ulp
    n = 300
    d = 10.0 ** np.linspace(-305, 305, n)
    a = cm(np.diag(d))
    lapack.dpotrf(a, n)
    got = mat(a, n, n)
    assert np.allclose(np.diag(got), np.sqrt(d), rtol=4 * EPS, atol=0)
    assert np.all(np.tril(got, -1) == 0.0)


@pytest.mark.parametrize("n, bad", [(200, 150), (900, 500), (900, 899)])
def test_dpotrf_not_spd_in_a_later_block(n, bad):
    # a failing pivot beyond the first diagonal block, in a later look-ahead block, and the last
    a0 = spd_gg(n)
    A = mat(a0, n, n)
    A[bad, bad] = -1e6
    with pytest.raises(al.LapackError) as e:
        lapack.dpotrf(cm(A), n)
    assert e.value.kind == al.SINGULAR


def test_dgetrf_singular_in_a_later_panel():
    n = 300
    A = mat(uniform(n * n), n, n)
    A[:, 200] = A[:, 10] * 2.0  # exactly dependent column beyond the first panels
    A[:, 200] = 0.0
    with pytest.raises(al.LapackError) as e:
        lapack.dgetrf(cm(A), n)
    assert e.value.kind == al.SINGULAR
