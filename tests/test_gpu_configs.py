"""GPU tests at the full sizes of the benchmark configurations.

The O(n^3) oracle cannot run at these sizes, so each result is checked through properties that
do not depend on size, plus the parts of the factorization the oracle can still compute:
  * GEMM (m = n = k in {4096, 8192}, all four transpose cases): 64 sampled rows and 64 sampled
    columns of C against numpy FP64 products of the same rows and columns, under the elementwise
    bound 4 k eps (|op(A)| |op(B)|)_ij;
  * LU (dgetrf n = 8192, dgetrf_rect 16384 x 2048): three Freivalds probes
    |P A x - L (U x)|_inf <= 4 n eps |A|_inf |x|_inf (c = 4), and the first 64-column panel
    against the oracle's dgetrf_rect of the same columns (the crate's arithmetic, so its pivots
    agree and its columns, moved by the later interchanges, agree bit for bit);
  * Cholesky (n = 8192 on G G^T + n I built with the library's dsyrk, upper triangle
    NaN-poisoned): Freivalds probes of A - L L^T, and the leading 64 x 64 block bit for bit
    against the oracle;
  * mixed precision (A = U diag(sigma) V^T, Haar U, V, sigma geometric 1..1/kappa):
    host-computed compensated backward error <= 1e-14 for every right-hand side, no fallback
    where kappa * 2^-24 < 1 (kappa <= 1e5), the FP64 fallback at kappa 1e8 / 1e11;
  * batched: the full 20,000-system launch, every info 0, backward error on 400 sampled systems;
  * served: orders 3000-4096, kappa up to 1e8 in FP32 mode (the in-server fallback), resident
    factors reused.
"""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
al = pytest.importorskip("algebra_linear_amd")
from algebra_linear_amd import _lib, lapack, serve  # noqa: E402

LIB = _lib.load()
EPS = np.finfo(float).eps


def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def uniform(n, g):
    return torch.rand(n, dtype=torch.float64, device=dev(), generator=g) * 2 - 1


# ------------------------------------------------------------------------------------------
# dgemm nn/tn/nt/tt at 4096 and 8192
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [4096, 8192])
def test_dgemm_all_variants_at_config_sizes(n):
    g = torch.Generator(device=dev())
    g.manual_seed(n)
    A, B = uniform(n * n, g), uniform(n * n, g)
    C = torch.zeros(n * n, dtype=torch.float64, device=dev())
    a = A.cpu().numpy().reshape(n, n).T     # column-major storage: a[i, j] = element (i, j)
    b = B.cpu().numpy().reshape(n, n).T
    rng = np.random.default_rng(n)
    rows = np.sort(rng.choice(n, 64, replace=False))
    cols = np.sort(rng.choice(n, 64, replace=False))
    for ta, tb in ((0, 0), (1, 0), (0, 1), (1, 1)):
        st = LIB.al_dev_dgemm(ta, tb, n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0, C.data_ptr(), n, None)
        assert st == 0, _lib.last_error()
        torch.cuda.synchronize()
        c = C.cpu().numpy().reshape(n, n).T
        opa = a.T if ta else a
        opb = b.T if tb else b
        for sel_r, sel_c in ((rows, slice(None)), (slice(None), cols)):
            want = opa[sel_r, :] @ opb[:, sel_c]
            bound = 4 * n * EPS * (np.abs(opa[sel_r, :]) @ np.abs(opb[:, sel_c]))
            assert np.all(np.abs(c[sel_r, sel_c] - want) <= bound), (ta, tb)


# ------------------------------------------------------------------------------------------
# dgetrf 8192, dgetrf_rect 16384 x 2048
# ------------------------------------------------------------------------------------------

def apply_pivots(rows, ipiv, k0, k1):
    """The crate's interchanges ipiv[k0..k1) applied in order to the rows of a 2-D array."""
    out = rows.copy()
    for k in range(k0, k1):
        p = int(ipiv[k])
        if p != k:
            out[[k, p]] = out[[p, k]]
    return out


@pytest.mark.parametrize("m,n", [(8192, 8192), (16384, 2048)])
def test_dgetrf_at_config_sizes(m, n):
    g = torch.Generator(device=dev())
    g.manual_seed(m + n)
    A0 = uniform(m * n, g)
    A = A0.clone()
    ipiv = torch.empty(n, dtype=torch.int32, device=dev())
    info = torch.zeros(1, dtype=torch.int32, device=dev())
    assert LIB.al_dev_dgetrf(m, n, A.data_ptr(), m, ipiv.data_ptr(), info.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    a0 = A0.cpu().numpy().reshape(n, m).T      # m x n
    lu = A.cpu().numpy().reshape(n, m).T
    piv = ipiv.cpu().numpy().astype(np.int64)
    assert np.all((piv >= np.arange(n)) & (piv < m))
    L = np.tril(lu, -1)
    L[np.arange(n), np.arange(n)] = 1.0
    U = np.triu(lu[:n, :])
    pa = apply_pivots(a0, piv, 0, n)
    anorm = np.abs(a0).sum(axis=1).max()
    rng = np.random.default_rng(1)
    for _ in range(3):                          # Freivalds probes
        x = rng.uniform(-1, 1, n)
        r = pa @ x - L @ (U @ x)
        assert np.abs(r).max() <= 4 * n * EPS * anorm * np.abs(x).max()
    # the first panel: the oracle's dgetrf_rect of the same 64 columns
    st, panel, ppiv = O.dgetrf_rect(np.ascontiguousarray(a0[:, :64].T).ravel(), m, 64)
    assert st == 0
    assert np.array_equal(ppiv, piv[:64])
    want = apply_pivots(panel.reshape(64, m).T, piv, 64, n)
    assert np.array_equal(lu[:, :64], want)


def test_dpotrf_at_config_size():
    n = 8192
    g = torch.Generator(device=dev())
    g.manual_seed(3)
    G = uniform(n * n, g)
    S = torch.zeros(n * n, dtype=torch.float64, device=dev())
    # A = G G^T + n I with the library's dsyrk (lower), then the upper triangle poisoned
    assert LIB.al_dev_dsyrk(0, n, n, 1.0, G.data_ptr(), n, 0.0, S.data_ptr(), n, None) == 0
    Sm = S.view(n, n)                           # Sm[j, i] = A[i, j]
    Sm.diagonal().add_(float(n))
    up = torch.tril(torch.ones(n, n, dtype=torch.bool, device=dev()), -1)   # Sm[j, i], j > i: A upper
    Sm[up] = float("nan")
    a_low = Sm.T.cpu().numpy()                  # A with NaN above the diagonal
    info = torch.zeros(1, dtype=torch.int32, device=dev())
    assert LIB.al_dev_dpotrf(n, S.data_ptr(), n, info.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    lf = Sm.T.cpu().numpy()
    assert np.all(np.isnan(lf[np.triu_indices(n, 1)]))    # the upper triangle was not written
    L = np.tril(lf)
    A = np.tril(a_low) + np.tril(a_low, -1).T
    anorm = np.abs(A).sum(axis=1).max()
    rng = np.random.default_rng(2)
    for _ in range(3):
        x = rng.uniform(-1, 1, n)
        r = A @ x - L @ (L.T @ x)
        assert np.abs(r).max() <= 4 * n * EPS * anorm * np.abs(x).max()
    # the leading 128 x 128 block is factored with the crate's own arithmetic
    st, l128 = O.dpotrf(np.ascontiguousarray(a_low[:128, :128].T).ravel(), 128)
    assert st == 0
    got = lf[:128, :128].T.ravel()              # column-major
    low = np.tril(np.ones((128, 128), bool)).T.ravel()
    assert np.array_equal(got[low], l128[low])


# ------------------------------------------------------------------------------------------
# mixed precision + FP64 refinement
# ------------------------------------------------------------------------------------------

_HAAR = {}


def haar_pair(n):
    if n not in _HAAR:
        _HAAR.clear()
        torch.cuda.empty_cache()
        g = torch.Generator(device=dev())
        g.manual_seed(1988 + n)

        def haar():
            q, r = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev(), generator=g))
            return q * torch.sign(torch.diagonal(r))
        _HAAR[n] = (haar(), haar())
    return _HAAR[n]


def config3(n, kappa):
    U, V = haar_pair(n)
    sig = kappa ** (-torch.arange(n, dtype=torch.float64, device=dev()) / max(n - 1, 1))
    return ((U * sig) @ V.T).T.contiguous()      # row-major A^T = column-major A


def solve_mixed(At, n, nrhs, factor, seed):
    g = torch.Generator(device=dev())
    g.manual_seed(seed)
    Bt = torch.randn((nrhs, n), dtype=torch.float64, device=dev(), generator=g)
    X = torch.zeros_like(Bt)
    stats = _lib.IrStats()
    opt = lapack.ir_options()
    st = LIB.al_dev_dgesv_mixed(n, nrhs, At.data_ptr(), n, Bt.data_ptr(), n, X.data_ptr(), n, factor,
                                ctypes.byref(opt), ctypes.byref(stats), None)
    torch.cuda.synchronize()
    return st, stats, Bt.cpu().numpy().ravel(), X.cpu().numpy().ravel()


CONFIG3 = [(4096, k, r) for k in (1e2, 1e5, 1e8, 1e11) for r in (1, 16)] + \
          [(16384, 1e2, 1), (16384, 1e2, 16), (16384, 1e11, 1)]


@pytest.mark.parametrize("n,kappa,nrhs", CONFIG3)
def test_mixed_fp32_at_config_sizes(n, kappa, nrhs):
    At = config3(n, kappa)
    st, stats, b, x = solve_mixed(At, n, nrhs, _lib.AL_FACTOR_FP32, int(kappa) % 1000 + nrhs)
    assert st == 0, (st, _lib.last_error())
    a = At.cpu().numpy().ravel()
    del At
    print(f"n={n} kappa={kappa:g} nrhs={nrhs}: steps {stats.iterations}, fell_back {stats.fell_back}, "
          f"berr {stats.backward_error:.2e}", flush=True)
    if kappa <= 1e5:
        assert stats.fell_back == 0 and stats.converged == 1
    if kappa >= 1e8:
        # kappa * 2^-24 > 1: the FP64 LU in the FP32 factor's row order produces the result
        assert stats.fell_back == 1
    for k in range(nrhs):
        assert O.backward_error(a, x[k * n:(k + 1) * n], b[k * n:(k + 1) * n], n) <= 1e-14


@pytest.mark.parametrize("n,kappa,nrhs", [(4096, 1e2, 1), (4096, 1e2, 16), (16384, 1e2, 1)])
def test_mixed_bf16_at_config_sizes(n, kappa, nrhs):
    # the bf16 form (split-bf16 trailing updates on the bf16 matrix cores) refines to FP64
    # accuracy without the fallback at kappa 1e2
    At = config3(n, kappa)
    st, stats, b, x = solve_mixed(At, n, nrhs, _lib.AL_FACTOR_BF16, 77 + nrhs)
    assert st == 0, (st, _lib.last_error())
    a = At.cpu().numpy().ravel()
    del At
    print(f"bf16 n={n} kappa={kappa:g} nrhs={nrhs}: steps {stats.iterations}, fell_back {stats.fell_back}, "
          f"berr {stats.backward_error:.2e}", flush=True)
    assert stats.fell_back == 0 and stats.converged == 1
    for k in range(nrhs):
        assert O.backward_error(a, x[k * n:(k + 1) * n], b[k * n:(k + 1) * n], n) <= 1e-14


# ------------------------------------------------------------------------------------------
# batched: the full 20,000-system launch
# ------------------------------------------------------------------------------------------

def batch_systems(count, seed):
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(np.log(8), np.log(512), count))).astype(np.int64)
    kinds = np.where(np.arange(count) % 2 == 0, _lib.AL_BATCH_GESV, _lib.AL_BATCH_POSV)
    nrhs = rng.integers(1, 9, count)
    order = np.argsort(-ns, kind="stable")
    return ns[order], kinds[order], nrhs[order]


@pytest.mark.parametrize("factor", [_lib.AL_FACTOR_FP64, _lib.AL_FACTOR_FP32])
def test_batched_twenty_thousand_systems(factor):
    count = 20000
    ns, kinds, nrhs = batch_systems(count, 1988)
    a_off = np.concatenate([[0], np.cumsum(ns * ns)[:-1]])
    b_off = np.concatenate([[0], np.cumsum(ns * nrhs)[:-1]])
    a_len, b_len = int((ns * ns).sum()), int((ns * nrhs).sum())
    g = torch.Generator(device=dev())
    g.manual_seed(1989)
    A = uniform(a_len, g)
    B = torch.randn(b_len, dtype=torch.float64, device=dev(), generator=g)
    for i in range(count):
        n, o = int(ns[i]), int(a_off[i])
        blk = A[o:o + n * n].view(n, n)
        if kinds[i] == _lib.AL_BATCH_POSV:
            blk.copy_(blk @ blk.T + n * torch.eye(n, dtype=torch.float64, device=dev()))
        else:
            blk.diagonal().add_(0.5 * n)
    a_host, b_host = A.cpu().numpy(), B.cpu().numpy()
    ent = (_lib.BatchEntry * count)(*[_lib.BatchEntry(int(a_off[i]), int(b_off[i]), int(ns[i]), int(nrhs[i]),
                                                      int(kinds[i]), 0) for i in range(count)])
    dE = torch.from_numpy(np.frombuffer(bytes(ent), dtype=np.uint8).copy()).to(dev())
    info = torch.full((count,), 7, dtype=torch.int32, device=dev())
    steps = torch.zeros(count, dtype=torch.int32, device=dev())
    st = LIB.al_dev_batch_solve(count, dE.data_ptr(), A.data_ptr(), a_len, B.data_ptr(), info.data_ptr(),
                                steps.data_ptr(), factor, None)
    assert st == 0, _lib.last_error()
    torch.cuda.synchronize()
    assert int(info.abs().max().item()) == 0
    x_host = B.cpu().numpy()
    for i in range(0, count, count // 400):
        n, r, ao, bo = int(ns[i]), int(nrhs[i]), int(a_off[i]), int(b_off[i])
        a = a_host[ao:ao + n * n]
        if kinds[i] == _lib.AL_BATCH_POSV:  # built symmetric, so the full matrix is the input
            pass
        for k in range(r):
            assert O.backward_error(a, x_host[bo + k * n:bo + (k + 1) * n], b_host[bo + k * n:bo + (k + 1) * n],
                                    n) <= 1e-14, (i, n, k)


# ------------------------------------------------------------------------------------------
# served: large orders, kappa up to 1e8, resident factors
# ------------------------------------------------------------------------------------------

def test_served_large_orders_and_fallback():
    specs = [(3000, 1e2, 3), (4096, 1e8, 2), (3500, 1e5, 8), (4096, 1e2, 1)]
    systems = []
    for i, (n, kappa, r) in enumerate(specs):
        At = config3(n, kappa)
        a = At.cpu().numpy().ravel()
        b = np.random.default_rng(i).standard_normal(n * r)
        systems.append((n, kappa, r, a, b))
        del At
    _HAAR.clear()
    torch.cuda.empty_cache()
    with serve.Server(slots=4) as srv:
        tickets = [(srv.submit(a, b, n, r, serve.GESV, f, keep=(f == serve.FP32 and kappa < 1e8)), n, kappa, r, a, b, f)
                   for (n, kappa, r, a, b) in systems for f in (serve.FP32, serve.FP64)]
        kept = []
        for t, n, kappa, r, a, b, f in tickets:
            x, res = srv.wait(t)
            assert res["status"] == 0, res
            if f == serve.FP32 and kappa >= 1e8:
                assert res["steps"] == -1      # refinement cannot converge: the FP64 fallback
            for k in range(r):
                assert O.backward_error(a, x[k * n:(k + 1) * n], b[k * n:(k + 1) * n], n) <= 1e-14
            if res["factor_id"] >= 0:
                kept.append((res["factor_id"], n, a))
        assert kept
        for fid, n, a in kept:                  # 30% of the served requests reuse a factor
            b2 = np.random.default_rng(fid).standard_normal(n * 2)
            x2, r2 = srv.wait(srv.submit(None, b2, n, 2, serve.GESV, serve.FP32, factor_id=fid))
            assert r2["status"] == 0
            for k in range(2):
                assert O.backward_error(a, x2[k * n:(k + 1) * n], b2[k * n:(k + 1) * n], n) <= 1e-14
            srv.release(fid)
