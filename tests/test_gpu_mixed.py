"""GPU tests of the mixed-precision solve (FP32 or bf16-update LU with tournament pivoting + FP64 refinement).

The crate has no mixed-precision solver, so the bar is the contract of its FP64 dgesv
measured as the normwise backward error
||b - Ax||_inf / (||A||_inf ||x||_inf + ||b||_inf) of every returned column, recomputed on
the CPU in FP64 with compensated sums (the oracle's or_backward_error), is <= 1e-14.
Matrices with a prescribed condition number come from the oracle's dlatms (the crate's
generator, src/lapack/families.rs:903-1079: A = U diag(sigma) V^T, sigma geometric from 1
to 1/kappa, U and V products of Householder reflectors from Normal vectors).
"""
import numpy as np
import pytest

import oracle as O
from matrices import cm, mat, singular3

pytestmark = pytest.mark.gpu

al = pytest.importorskip("algebra_linear_amd")
lapack = al.lapack
LapackError = al.errors.LapackError
SEEDY = [1988, 1989, 1990, 1991]
TOL = 1e-14


def uniform(count, seed=SEEDY):
    return O.larnv(O.SYMMETRIC, seed, count)[0]


def normal(count, seed):
    return O.larnv(O.NORMAL, seed, count)[0]


def conditioned(n, kappa, seed=SEEDY):
    info, a, _, _ = O.dlatms(O.NORMAL, O.NONSYMMETRIC, 3, kappa, 1.0, O.FULL, n, n, seed)
    assert info == 0
    return a


def solve(a, b, n, nrhs, factor, **kw):
    x = b.copy()
    a_before = np.array(a, copy=True)
    st = lapack.dgesv_mixed(a, x, n, nrhs, factor, lapack.ir_options(**kw))
    np.testing.assert_array_equal(a, a_before)  # A is never modified
    return x, st


@pytest.mark.parametrize("n", [1, 3, 64, 100, 257, 512])
def test_fp32_uniform_converges_without_fallback(n):
    a = uniform(n * n)
    b = uniform(n, [7, 11, 13, 17])
    x, st = solve(a, b, n, 1, "fp32")
    berr = O.backward_error(a, x, b, n)
    assert st["fell_back"] == 0 and st["converged"] == 1 and st["factor_breakdown"] == 0
    assert berr <= TOL, (berr, st)
    # the library's own (GPU, FP64) estimate agrees with the CPU's compensated one
    assert st["backward_error"] <= TOL
    assert st["iterations"] <= 10


@pytest.mark.parametrize("kappa", [1e2, 1e5])
@pytest.mark.parametrize("n", [256, 320])
def test_fp32_conditioned_converges_without_fallback(n, kappa):
    a = conditioned(n, kappa)
    b = normal(n, [3, 5, 7, 9])
    x, st = solve(a, b, n, 1, "fp32")
    assert st["fell_back"] == 0 and st["converged"] == 1, st
    assert O.backward_error(a, x, b, n) <= TOL


@pytest.mark.parametrize("kappa", [1e8, 1e11])
def test_ill_conditioned_reaches_fp64_accuracy_with_fallback_if_needed(kappa):
    n = 256
    a = conditioned(n, kappa)
    b = normal(n, [3, 5, 7, 9])
    x, st = solve(a, b, n, 1, "fp32")
    assert st["converged"] == 1, st
    assert O.backward_error(a, x, b, n) <= TOL
    if kappa >= 1e11:
        # kappa * u_fp32 >> 1: FP32 refinement cannot converge, and the FP64 LU in the FP32
        # factor's row order (no pivot search) with FP64 refinement produces the result
        assert st["fell_back"] == 1


def test_fp32_breakdown_falls_back_to_the_pivoted_fp64_lu():
    # rows 0 and 1 differ only below FP32's resolution: the FP32 factor meets an exact zero
    # pivot, its row order is incomplete, and the partially pivoted FP64 LU solves the system
    n = 130
    A = np.eye(n)
    A[0, 0] = A[0, 1] = A[1, 0] = 1.0
    A[1, 1] = 1.0 + 2.0 ** -30
    A[5:, :5] = 0.25
    a = cm(A)
    b = normal(n, [3, 5, 7, 9])
    x, st = solve(a, b, n, 1, "fp32")
    assert st["factor_breakdown"] > 0 and st["fell_back"] == 2 and st["converged"] == 1, st
    assert O.backward_error(a, x, b, n) <= TOL


def test_refinement_failure_without_fallback_raises_convergence():
    n = 256
    a = conditioned(n, 1e11)
    b = normal(n, [3, 5, 7, 9])
    x = b.copy()
    with pytest.raises(LapackError) as e:
        lapack.dgesv_mixed(a, x, n, 1, "fp32", lapack.ir_options(allow_fallback=False))
    assert e.value.kind == "Convergence"


@pytest.mark.parametrize("n,kappa", [(256, 1e2), (600, 10.0), (1100, 1e2), (1500, 1e3)])
def test_bf16_factor_reaches_fp64_accuracy(n, kappa):
    # n = 1100 pads to 1152: two 512-wide outer blocks plus a 128-wide remainder, so the
    # bf16 trailing update runs on full and edge tiles.  The result must come from the bf16
    # factorization and FP64 refinement, not from the FP64 fallback.
    a = conditioned(n, kappa)
    b = normal(n, [3, 5, 7, 9])
    x, st = solve(a, b, n, 1, "bf16")
    assert st["converged"] == 1 and st["fell_back"] == 0, st
    assert O.backward_error(a, x, b, n) <= TOL


@pytest.mark.parametrize("nrhs", [4, 6, 16, 20])
def test_multiple_right_hand_sides(nrhs):
    n = 300
    a = uniform(n * n)
    b = normal(n * nrhs, [21, 22, 23, 25])
    x, st = solve(a, b, n, nrhs, "fp32")
    assert st["fell_back"] == 0 and st["converged"] == 1
    for j in range(nrhs):
        assert O.backward_error(a, x[j * n:(j + 1) * n], b[j * n:(j + 1) * n], n) <= TOL
    # each column agrees with its single-right-hand-side solve: the factor and the refinement
    # arithmetic do not depend on the other columns, but a column solved beside others may take
    # more refinement steps (the group stops when all its columns are done), so the two agree
    # to the solutions' accuracy (forward error <= cond(A) * backward error), not bit for bit
    for j in (0, nrhs - 1):
        xj, stj = solve(a, b[j * n:(j + 1) * n].copy(), n, 1, "fp32")
        assert O.backward_error(a, xj, b[j * n:(j + 1) * n], n) <= TOL
        col = x[j * n:(j + 1) * n]
        assert np.abs(xj - col).max() <= 1e-10 * np.abs(col).max()


def test_fp64_factor_is_the_plain_pivoted_solve():
    n = 200
    a = uniform(n * n)
    b = uniform(n * 3, [9, 9, 9, 9])
    x, st = solve(a, b, n, 3, "fp64")
    xref = b.copy()
    aref = a.copy()
    lapack.dgesv_nrhs(aref, xref, n, 3)
    np.testing.assert_array_equal(x, xref)
    assert st["fell_back"] == 0 and st["iterations"] == 0 and st["converged"] == 1


def test_inconsistent_singular_system_raises_singular():
    # the crate's singular matrix (src/lapack/mod.rs:497-508) with b outside its range: no
    # iterate reaches a small backward error, and the FP64 fallback meets the crate's exactly
    # zero pivot, so the call reports Singular as lapack::dgesv does
    a = singular3()
    b = np.array([1.0, 0.0, 0.0])
    with pytest.raises(LapackError) as e:
        lapack.dgesv_mixed(a, b.copy(), 3, 1, "fp32")
    assert e.value.kind == "Singular"


def test_consistent_singular_system_is_singular_or_solved_backward_stably():
    # b in the range of the singular matrix: either the FP32 elimination meets an exactly zero
    # pivot (then the FP64 fallback reports Singular, as lapack::dgesv), or refinement returns
    # one of the solutions with an FP64-level backward error
    a = singular3()
    b = mat(a, 3, 3) @ np.array([1.0, 1.0, 1.0])
    x = b.copy()
    try:
        st = lapack.dgesv_mixed(a, x, 3, 1, "fp32")
    except LapackError as e:
        assert e.kind == "Singular"
        return
    assert st["converged"] == 1
    assert O.backward_error(a, x, b, 3) <= TOL


def test_bad_shapes_and_empty():
    with pytest.raises(LapackError) as e:
        lapack.dgesv_mixed(np.zeros(8), np.zeros(3), 3)
    assert e.value.kind == "BadShape"
    st = lapack.dgesv_mixed(np.zeros(0), np.zeros(0), 0)
    assert st["converged"] == 1


def test_results_are_reproducible():
    n = 384
    a = conditioned(n, 1e3)
    b = normal(n, [3, 5, 7, 9])
    x1, s1 = solve(a, b, n, 1, "fp32")
    x2, s2 = solve(a, b, n, 1, "fp32")
    np.testing.assert_array_equal(x1, x2)
    assert s1 == s2


def test_deep_tournament_tree_converges_and_is_reproducible():
    # n = 4160: the first 64-column panel has 4160 rows, i.e. 17 leaves of 256 and four tree
    # levels (17 -> 5 -> 2 -> 1), so nominees pass between workgroups (and XCDs) three times;
    # a nominee read before it is visible would pick wrong pivot rows (no convergence) or
    # change the result from run to run
    n = 4160
    a = uniform(n * n)
    b = uniform(n, [7, 11, 13, 17])
    x1, s1 = solve(a, b, n, 1, "fp32")
    x2, s2 = solve(a, b, n, 1, "fp32")
    assert s1["fell_back"] == 0 and s1["converged"] == 1 and s1["factor_breakdown"] == 0, s1
    assert O.backward_error(a, x1, b, n) <= TOL
    np.testing.assert_array_equal(x1, x2)
    assert s1 == s2


def test_device_api_on_a_torch_stream_with_padded_leading_dimensions():
    torch = pytest.importorskip("torch")
    import ctypes
    from algebra_linear_amd import _lib
    n, nrhs, lda, ldb = 190, 2, 200, 197
    a = mat(uniform(n * n), n, n)
    b = mat(normal(n * nrhs, [4, 4, 4, 4]), n, nrhs)
    dA = torch.zeros(n * lda, dtype=torch.float64, device="cuda")
    dB = torch.zeros(nrhs * ldb, dtype=torch.float64, device="cuda")
    dX = torch.full((nrhs * ldb,), 7.0, dtype=torch.float64, device="cuda")
    for j in range(n):
        dA[j * lda:j * lda + n] = torch.from_numpy(a[:, j].copy())
    for j in range(nrhs):
        dB[j * ldb:j * ldb + n] = torch.from_numpy(b[:, j].copy())
    stream = torch.cuda.Stream()
    st = _lib.IrStats()
    opt = lapack.ir_options()
    with torch.cuda.stream(stream):
        rc = _lib.load().al_dev_dgesv_mixed(n, nrhs, dA.data_ptr(), lda, dB.data_ptr(), ldb, dX.data_ptr(), ldb,
                                            _lib.AL_FACTOR_FP32, ctypes.byref(opt), ctypes.byref(st),
                                            ctypes.c_void_p(stream.cuda_stream))
    assert rc == 0, _lib.last_error()
    stream.synchronize()
    xh = dX.cpu().numpy()
    for j in range(nrhs):
        xj = xh[j * ldb:j * ldb + n]
        assert O.backward_error(cm(a), xj, b[:, j].copy(), n) <= TOL
        assert np.all(xh[j * ldb + n:(j + 1) * ldb] == 7.0)  # rows beyond n untouched


def test_first_call_on_a_fresh_thread_refines_without_fallback():
    # a new thread gets a fresh stream and workspace; the first solve on them must see
    # initialized device-side flags (a regression test: the flag arrays were cleared on the
    # null stream, which does not order the library's non-blocking stream)
    import threading
    n = 600
    a = conditioned(n, 1e2)
    b = normal(n, [3, 5, 7, 9])
    out = {}

    def run():
        x = b.copy()
        out["st"] = lapack.dgesv_mixed(a, x, n, 1, "fp32")
        out["x"] = x

    th = threading.Thread(target=run)
    th.start()
    th.join()
    assert out["st"]["fell_back"] == 0 and out["st"]["converged"] == 1, out["st"]
    assert O.backward_error(a, out["x"], b, n) <= TOL
