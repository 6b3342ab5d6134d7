"""GPU tests of the continuous-batching solve server.

Each request's solution is checked against the crate's contract (normwise backward error of
every column <= 1e-14, computed on the CPU in FP64 with compensated sums), and each result is
shown to be the same -- bit for bit -- when the request is served alone and in a concurrent
mix with others.
"""
import numpy as np
import pytest

import oracle as O
from matrices import singular3

pytestmark = pytest.mark.gpu

al = pytest.importorskip("algebra_linear_amd")
from algebra_linear_amd import serve  # noqa: E402

TOL = 1e-14


def system(n, nrhs, kind, seed, kappa=None):
    rng = np.random.default_rng(seed)
    if kind == serve.POSV:
        g = rng.uniform(-1, 1, (n, n))
        a = g @ g.T + n * np.eye(n)
    elif kappa is not None:
        info, a, _, _ = O.dlatms(O.NORMAL, O.NONSYMMETRIC, 3, kappa, 1.0, O.FULL, n, n, [seed % 4096, 7, 11, 13])
        assert info == 0
        return a, rng.standard_normal(n * nrhs)
    else:
        a = rng.uniform(-1, 1, (n, n))
    return a.reshape(-1, order="F"), rng.standard_normal(n * nrhs)


def sym_full(a, n):
    a2 = a.reshape(n, n, order="F")
    return (np.tril(a2) + np.tril(a2, -1).T).reshape(-1, order="F")


def check(a, x, b, n, nrhs, kind):
    aa = sym_full(a, n) if kind == serve.POSV else a
    for k in range(nrhs):
        assert O.backward_error(aa, x[k * n:(k + 1) * n], b[k * n:(k + 1) * n], n) <= TOL


MIX = [(n, nrhs, kind, factor)
       for i, (n, nrhs) in enumerate([(16, 1), (100, 3), (300, 8), (600, 2), (1000, 1), (40, 5), (2000, 1),
                                      (513, 2), (64, 8), (1500, 4)])
       for kind in (serve.GESV, serve.POSV)
       for factor in (serve.FP64, serve.FP32)]


def test_mixed_requests_served_concurrently_match_alone_bit_for_bit():
    systems = [system(n, r, k, 100 + i, kappa=1e3 if (k == serve.GESV and i % 3 == 0) else None)
               for i, (n, r, k, f) in enumerate(MIX)]
    # alone: one slot, one request at a time
    alone = []
    with serve.Server(slots=1) as srv:
        for (n, r, k, f), (a, b) in zip(MIX, systems):
            x, res = srv.wait(srv.submit(a, b, n, r, k, f))
            assert res["status"] == 0, (n, k, f, res)
            alone.append(x.copy())
    # in the mix: every request submitted at once to four slots
    with serve.Server(slots=4) as srv:
        tickets = [srv.submit(a, b, n, r, k, f) for (n, r, k, f), (a, b) in zip(MIX, systems)]
        results = [srv.wait(t) for t in tickets]
    for (n, r, k, f), (a, b), xa, (x, res) in zip(MIX, systems, alone, results):
        assert res["status"] == 0
        check(a, x, b, n, r, k)
        np.testing.assert_array_equal(x, xa)


def test_small_requests_waiting_together_share_one_launch():
    with serve.Server(slots=1) as srv:
        big = system(2500, 1, serve.GESV, 77)
        t_big = srv.submit(big[0], big[1], 2500)  # keeps the only slot busy
        small = [system(n, 2, serve.GESV if i % 2 else serve.POSV, 200 + i) for i, n in enumerate(range(20, 60, 4))]
        kinds = [serve.GESV if i % 2 else serve.POSV for i in range(len(small))]
        ts = [srv.submit(a, b, len(b) // 2, 2, k) for (a, b), k in zip(small, kinds)]
        srv.wait(t_big)
        sizes = []
        for (a, b), k, t in zip(small, kinds, ts):
            x, res = srv.wait(t)
            assert res["status"] == 0
            check(a, x, b, len(b) // 2, 2, k)
            sizes.append(res["batch_size"])
        assert max(sizes) > 1, sizes


@pytest.mark.parametrize("factor", [serve.FP64, serve.FP32])
@pytest.mark.parametrize("kind", [serve.GESV, serve.POSV])
def test_resident_factor_serves_new_right_hand_sides(factor, kind):
    n = 700
    a, b = system(n, 2, kind, 5)
    with serve.Server(slots=2) as srv:
        x, res = srv.wait(srv.submit(a, b, n, 2, kind, factor, keep=True))
        assert res["status"] == 0 and res["factor_id"] >= 0
        check(a, x, b, n, 2, kind)
        fid = res["factor_id"]
        rng = np.random.default_rng(9)
        for _ in range(3):
            b2 = rng.standard_normal(n * 3)
            x2, r2 = srv.wait(srv.submit(None, b2, n, 3, kind, factor, factor_id=fid))
            assert r2["status"] == 0
            check(a, x2, b2, n, 3, kind)
        srv.release(fid)
        _, r3 = srv.wait(srv.submit(None, b[:n], n, 1, kind, factor, factor_id=fid))
        assert r3["status"] < 0  # released: unknown factor


@pytest.mark.parametrize("factor", [serve.FP64, serve.FP32])
def test_released_factor_blocks_serve_later_keeps(factor):
    # keep, release, keep a slightly smaller system (its block comes from the released one),
    # keep a much smaller one (a fresh block): every factor solves its own system, and the
    # reused block holds nothing of the earlier factor
    with serve.Server(slots=1) as srv:
        ids = []
        for n, seed in ((600, 21), (560, 22), (200, 23)):
            a, b = system(n, 1, serve.GESV, seed)
            x, res = srv.wait(srv.submit(a, b, n, 1, serve.GESV, factor, keep=True))
            assert res["status"] == 0 and res["factor_id"] >= 0
            check(a, x, b, n, 1, serve.GESV)
            rng = np.random.default_rng(seed)
            b2 = rng.standard_normal(n * 2)
            x2, r2 = srv.wait(srv.submit(None, b2, n, 2, serve.GESV, factor, factor_id=res["factor_id"]))
            assert r2["status"] == 0
            check(a, x2, b2, n, 2, serve.GESV)
            if n == 600:
                srv.release(res["factor_id"])
            else:
                ids.append(res["factor_id"])
        for fid in ids:
            srv.release(fid)


def test_requests_larger_than_the_budget_still_run_alone():
    with serve.Server(hbm_budget_bytes=1 << 20, slots=3) as srv:
        specs = [(800, 1), (20, 2), (900, 1)]
        sys_ = [system(n, r, serve.GESV, 40 + i) for i, (n, r) in enumerate(specs)]
        tickets = [srv.submit(a, b, n, r) for (n, r), (a, b) in zip(specs, sys_)]
        for (n, r), (a, b), t in zip(specs, sys_, tickets):
            x, res = srv.wait(t)
            assert res["status"] == 0
            check(a, x, b, n, r, serve.GESV)


def test_singular_request_does_not_disturb_its_batch():
    with serve.Server(slots=1) as srv:
        a0, b0 = system(30, 1, serve.GESV, 1)
        t0 = srv.submit(a0, b0, 30)
        t1 = srv.submit(singular3(), np.array([1.0, 0.0, 0.0]), 3)
        a2, b2 = system(50, 2, serve.GESV, 2)
        t2 = srv.submit(a2, b2, 50, 2)
        x0, r0 = srv.wait(t0)
        _, r1 = srv.wait(t1)
        x2, r2 = srv.wait(t2)
        assert r1["status"] == 1  # Singular
        assert r0["status"] == 0 and r2["status"] == 0
        check(a0, x0, b0, 30, 1, serve.GESV)
        check(a2, x2, b2, 50, 2, serve.GESV)
