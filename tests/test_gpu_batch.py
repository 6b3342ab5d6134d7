"""GPU tests of the batched solver (many independent dgesv / dposv systems in one launch).

Every system is checked on its own against the crate's contract: the normwise backward
error of each solution column, computed on the CPU in FP64 with compensated sums, is at most
1e-14; Cholesky factors satisfy |A - L L^T|_1 / (n eps |A|_1) < 30 (the bar of the crate's
pot03-style checks) and leave the upper triangle untouched; an exactly singular system reports
the crate's Singular as a positive status without disturbing its neighbours.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from matrices import singular3

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
al = pytest.importorskip("algebra_linear_amd")
from algebra_linear_amd import _lib  # noqa: E402

LIB = _lib.load()
TOL = 1e-14
EPS = np.finfo(float).eps


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def make_batch(specs, seed=11):
    """specs: list of (n, nrhs, kind).  Returns host pools, entry array and per-system views."""
    rng = np.random.default_rng(seed)
    a_parts, b_parts, entries = [], [], []
    a_off = b_off = 0
    for n, nrhs, kind in specs:
        if kind == _lib.AL_BATCH_POSV:
            g = rng.uniform(-1, 1, (n, n))
            a = g @ g.T + n * np.eye(n)
        else:
            a = rng.uniform(-1, 1, (n, n)) + 0.5 * n * np.eye(n) * (rng.uniform() < 0.5)
        b = rng.standard_normal((n, nrhs))
        a_parts.append(a.reshape(-1, order="F"))
        b_parts.append(b.reshape(-1, order="F"))
        entries.append(_lib.BatchEntry(a_off, b_off, n, nrhs, kind, 0))
        a_off += n * n
        b_off += n * nrhs
    A = np.concatenate(a_parts) if a_parts else np.zeros(0)
    B = np.concatenate(b_parts) if b_parts else np.zeros(0)
    return A, B, entries


def entries_to_device(entries, dev):
    arr = (_lib.BatchEntry * len(entries))(*entries)
    raw = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
    return torch.from_numpy(raw).to(dev)


def run_batch(specs, factor, seed=11):
    dev = _dev()
    A, B, entries = make_batch(specs, seed)
    dA = torch.from_numpy(A.copy()).to(dev)
    dB = torch.from_numpy(B.copy()).to(dev)
    dE = entries_to_device(entries, dev)
    info = torch.full((len(entries),), 99, dtype=torch.int32, device=dev)
    steps = torch.full((len(entries),), 99, dtype=torch.int32, device=dev)
    rc = LIB.al_dev_batch_solve(len(entries), dE.data_ptr(), dA.data_ptr(), A.size, dB.data_ptr(),
                                info.data_ptr(), steps.data_ptr(), factor, None)
    assert rc == 0, _lib.last_error()
    torch.cuda.synchronize()
    return A, B, entries, dA.cpu().numpy(), dB.cpu().numpy(), info.cpu().numpy(), steps.cpu().numpy()


def system(pool, off, n, m):
    return pool[off:off + n * m]


SPECS = [(n, nrhs, kind)
         for n, nrhs in [(1, 1), (2, 3), (7, 1), (16, 2), (17, 8), (33, 9), (64, 1), (65, 4), (100, 17),
                         (128, 1), (200, 3), (255, 1), (256, 2), (300, 1), (511, 2), (512, 1)]
         for kind in (_lib.AL_BATCH_GESV, _lib.AL_BATCH_POSV)]


@pytest.mark.parametrize("factor", [_lib.AL_FACTOR_FP64, _lib.AL_FACTOR_FP32])
def test_every_system_of_an_uneven_batch_is_solved(factor):
    A, B, entries, A_out, X, info, steps = run_batch(SPECS, factor)
    for i, e in enumerate(entries):
        n, nrhs = e.n, e.nrhs
        assert info[i] == 0, (i, n)
        a = system(A, e.a_off, n, n)
        if e.kind == _lib.AL_BATCH_POSV:
            a2 = a.reshape(n, n, order="F")
            a = np.tril(a2) + np.tril(a2, -1).T  # the solve uses the lower triangle only
            a = a.reshape(-1, order="F")
        for k in range(nrhs):
            x = system(X, e.b_off + k * n, n, 1)
            b = system(B, e.b_off + k * n, n, 1)
            assert O.backward_error(a, x, b, n) <= TOL, (i, n, k)
        if factor == _lib.AL_FACTOR_FP64:
            assert steps[i] == 0
        else:
            assert steps[i] >= 1 or steps[i] == -1
            if steps[i] >= 1:  # refined: A untouched
                np.testing.assert_array_equal(system(A_out, e.a_off, n, n), system(A, e.a_off, n, n))


def test_cholesky_factors_and_untouched_upper_triangle():
    specs = [(n, 1, _lib.AL_BATCH_POSV) for n in (5, 64, 129, 300)]
    A, B, entries, A_out, X, info, steps = run_batch(specs, _lib.AL_FACTOR_FP64)
    for i, e in enumerate(entries):
        n = e.n
        a = system(A, e.a_off, n, n).reshape(n, n, order="F")
        f = system(A_out, e.a_off, n, n).reshape(n, n, order="F")
        L = np.tril(f)
        res = np.abs(a - L @ L.T).sum(axis=0).max() / (n * EPS * np.abs(a).sum(axis=0).max())
        assert res < 30, (n, res)
        iu = np.triu_indices(n, 1)
        np.testing.assert_array_equal(f[iu], a[iu])
        # the oracle's Cholesky (the crate's algorithm) agrees to rounding
        st, l_ref = O.dpotrf(system(A, e.a_off, n, n), n)
        assert st == 0
        np.testing.assert_allclose(L, np.tril(l_ref.reshape(n, n, order="F")), rtol=1e-12, atol=1e-12 * n)


def test_lu_factors_reproduce_the_oracle_pivoting():
    # the batch LU chooses the crate's pivots (first maximum): the solution of a system with a
    # well separated pivot sequence matches the oracle's dgesv to rounding
    specs = [(n, 1, _lib.AL_BATCH_GESV) for n in (3, 31, 97)]
    A, B, entries, A_out, X, info, steps = run_batch(specs, _lib.AL_FACTOR_FP64, seed=5)
    for i, e in enumerate(entries):
        n = e.n
        st, _, x_ref = O.dgesv(system(A, e.a_off, n, n).copy(), system(B, e.b_off, n, 1).copy())
        assert st == 0
        np.testing.assert_allclose(system(X, e.b_off, n, 1), x_ref, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("factor", [_lib.AL_FACTOR_FP64, _lib.AL_FACTOR_FP32])
def test_singular_and_oversized_systems_report_status_without_disturbing_others(factor):
    dev = _dev()
    specs = [(40, 2, _lib.AL_BATCH_GESV), (3, 1, _lib.AL_BATCH_GESV), (50, 1, _lib.AL_BATCH_POSV),
             (30, 1, _lib.AL_BATCH_GESV)]
    A, B, entries = make_batch(specs, seed=3)
    # system 1: the crate's singular matrix (src/lapack/mod.rs:497-508)
    A[entries[1].a_off:entries[1].a_off + 9] = singular3()
    # system 2: not positive definite
    a2 = -np.eye(50)
    A[entries[2].a_off:entries[2].a_off + 2500] = a2.reshape(-1)
    dA = torch.from_numpy(A.copy()).to(dev)
    dB = torch.from_numpy(B.copy()).to(dev)
    dE = entries_to_device(entries, dev)
    info = torch.full((4,), 99, dtype=torch.int32, device=dev)
    rc = LIB.al_dev_batch_solve(4, dE.data_ptr(), dA.data_ptr(), A.size, dB.data_ptr(), info.data_ptr(), None,
                                factor, None)
    assert rc == 0
    torch.cuda.synchronize()
    inf = info.cpu().numpy()
    assert inf[1] > 0 and inf[2] == 1, inf
    assert inf[0] == 0 and inf[3] == 0
    X = dB.cpu().numpy()
    for i in (0, 3):
        e = entries[i]
        for k in range(e.nrhs):
            assert O.backward_error(system(A, e.a_off, e.n, e.n), system(X, e.b_off + k * e.n, e.n, 1),
                                    system(B, e.b_off + k * e.n, e.n, 1), e.n) <= TOL


def test_order_above_512_is_rejected_per_system():
    dev = _dev()
    specs = [(513, 1, _lib.AL_BATCH_GESV), (4, 1, _lib.AL_BATCH_GESV)]
    A, B, entries = make_batch(specs)
    dA, dB = torch.from_numpy(A).to(dev), torch.from_numpy(B.copy()).to(dev)
    dE = entries_to_device(entries, dev)
    info = torch.zeros(2, dtype=torch.int32, device=dev)
    assert LIB.al_dev_batch_solve(2, dE.data_ptr(), dA.data_ptr(), A.size, dB.data_ptr(), info.data_ptr(), None,
                                  _lib.AL_FACTOR_FP64, None) == 0
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [-4, 0]


def test_graph_captured_batch_replays_match_eager():
    dev = _dev()
    specs = [(n, 2, kind) for n in (300, 64, 17) for kind in (_lib.AL_BATCH_GESV, _lib.AL_BATCH_POSV)]
    A, B, entries = make_batch(specs, seed=9)
    A0 = torch.from_numpy(A).to(dev)
    B0 = torch.from_numpy(B).to(dev)
    dE = entries_to_device(entries, dev)
    dA, dB = A0.clone(), B0.clone()
    info = torch.zeros(len(entries), dtype=torch.int32, device=dev)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for factor in (_lib.AL_FACTOR_FP64, _lib.AL_FACTOR_FP32):  # warm-up: workspaces allocated
            dA.copy_(A0)
            dB.copy_(B0)
            assert LIB.al_dev_batch_solve(len(entries), dE.data_ptr(), dA.data_ptr(), A.size, dB.data_ptr(),
                                          info.data_ptr(), None, factor, ctypes.c_void_p(s.cuda_stream)) == 0
    s.synchronize()
    eager = dB.cpu().numpy().copy()  # FP32-mode result
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        dA.copy_(A0)
        dB.copy_(B0)
        LIB.al_dev_batch_solve(len(entries), dE.data_ptr(), dA.data_ptr(), A.size, dB.data_ptr(), info.data_ptr(),
                               None, _lib.AL_FACTOR_FP32, ctypes.c_void_p(s.cuda_stream))
    for _ in range(3):
        dB.fill_(0.0)  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dB.cpu().numpy(), eager)
        assert info.cpu().tolist() == [0] * len(entries)


def test_small_orders_are_solved_bit_for_bit_as_the_crate_solves_them():
    # orders <= 64 go to the one-wave-per-system kernel, which follows the crate's own loops:
    # dgetrf_with_pivots + dgetrs (src/lapack/mod.rs:112-163, :330-354), dpotrf + dtrsv +
    # dtrsv_transpose (:368-399; src/blas.rs:1355-1438), unfused and in the crate's order
    specs = [(n, nrhs, kind) for n, nrhs in [(1, 1), (2, 2), (5, 3), (16, 8), (31, 1), (33, 9), (48, 2), (64, 5)]
             for kind in (_lib.AL_BATCH_GESV, _lib.AL_BATCH_POSV)]
    A, B, entries, A_out, X, info, steps = run_batch(specs, _lib.AL_FACTOR_FP64, seed=21)
    for i, e in enumerate(entries):
        n, nrhs = e.n, e.nrhs
        assert info[i] == 0, (i, n)
        a = system(A, e.a_off, n, n)
        if e.kind == _lib.AL_BATCH_GESV:
            st, lu, ipiv, _ = O.dgetrf(a, n)
            assert st == 0
            np.testing.assert_array_equal(system(A_out, e.a_off, n, n), lu)
            for k in range(nrhs):
                want = O.dgetrs(lu, ipiv, system(B, e.b_off + k * n, n, 1))
                np.testing.assert_array_equal(system(X, e.b_off + k * n, n, 1), want)
        else:
            st, l = O.dpotrf(a, n)
            assert st == 0
            low = np.tril(np.ones((n, n), bool)).reshape(-1, order="F")
            np.testing.assert_array_equal(system(A_out, e.a_off, n, n)[low], l[low])
            for k in range(nrhs):
                y = O.dtrsv(l, system(B, e.b_off + k * n, n, 1), n, True, False)
                want = O.dtrsv_transpose(l, y, n, True, False)
                np.testing.assert_array_equal(system(X, e.b_off + k * n, n, 1), want)


def test_every_bucket_and_its_boundaries():
    # the bucket edges (64/65, 128/129, 256/257, 512) in one call, in the caller's order, mixed
    specs = [(n, 1 + n % 5, kind) for n in (512, 3, 257, 64, 256, 65, 129, 128, 1, 300, 100, 40)
             for kind in (_lib.AL_BATCH_GESV, _lib.AL_BATCH_POSV)]
    for factor in (_lib.AL_FACTOR_FP64, _lib.AL_FACTOR_FP32):
        A, B, entries, A_out, X, info, steps = run_batch(specs, factor, seed=31)
        for i, e in enumerate(entries):
            n = e.n
            assert info[i] == 0, (factor, i, n)
            a = system(A, e.a_off, n, n)
            if e.kind == _lib.AL_BATCH_POSV:
                a2 = a.reshape(n, n, order="F")
                a = (np.tril(a2) + np.tril(a2, -1).T).reshape(-1, order="F")
            for k in range(e.nrhs):
                assert O.backward_error(a, system(X, e.b_off + k * n, n, 1), system(B, e.b_off + k * n, n, 1),
                                        n) <= TOL, (factor, i, n, k)


def test_tuning_table_entries_change_the_launch_not_the_result():
    specs = [(n, 2, kind) for n in (70, 150, 300) for kind in (_lib.AL_BATCH_GESV, _lib.AL_BATCH_POSV)]
    kb, w = ctypes.c_int(), ctypes.c_int()
    saved = []
    for b in range(4):
        assert LIB.al_batch_tune_get(b, ctypes.byref(kb), ctypes.byref(w)) == 0
        saved.append((kb.value, w.value))
    try:
        _, _, _, _, X1, info1, _ = run_batch(specs, _lib.AL_FACTOR_FP64, seed=41)
        assert LIB.al_batch_tune_set(1, 32, 1) == 0 and LIB.al_batch_tune_set(2, 32, 1) == 0
        assert LIB.al_batch_tune_set(3, 32, 1) != 0          # bucket 3 holds a 16-wide panel only
        A, B, entries, _, X2, info2, _ = run_batch(specs, _lib.AL_FACTOR_FP64, seed=41)
        assert list(info2) == [0] * len(specs)
        for i, e in enumerate(entries):
            n = e.n
            a = system(A, e.a_off, n, n)
            if e.kind == _lib.AL_BATCH_POSV:
                a2 = a.reshape(n, n, order="F")
                a = (np.tril(a2) + np.tril(a2, -1).T).reshape(-1, order="F")
            for k in range(e.nrhs):
                assert O.backward_error(a, system(X2, e.b_off + k * n, n, 1), system(B, e.b_off + k * n, n, 1),
                                        n) <= TOL
    finally:
        for b, (k_, w_) in enumerate(saved):
            LIB.al_batch_tune_set(b, k_ if b else 16, w_)
