"""GPU tests of the error paths of the bounded device-side waits, forced from the host.

Every cooperative wait in the library (the FP64 panel's column exchange, the one-launch
triangular solves' block flags) gives up after a limit and raises an abort flag that ends every
waiting workgroup normally.  al_debug_set_wait_limit lowers that limit to one tick of the wall
clock, so the first wait that does not succeed at once gives up; no kernel faults or hangs.  The
tests check that each entry point then reports AL_ERR_DEVICE instead of a partial result, that
the server keeps no half-built factorization, and that the next call with the default limit is
correct again.  The last test runs FP64 factorizations on several server slots at once, whose
cooperative panels together need more workgroups than half the device's CUs (a panel may run on
one half of the chip), and checks that the process-wide reservation keeps them within that count
and every result correct.
"""
import ctypes
import threading
import time

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

al = pytest.importorskip("algebra_linear_amd")
from algebra_linear_amd import _lib, serve  # noqa: E402

L = _lib.load()
ERR_DEVICE = _lib.AL_ERR_DEVICE


@pytest.fixture
def tiny_wait_limit():
    L.al_debug_set_wait_limit(1)
    try:
        yield
    finally:
        L.al_debug_set_wait_limit(0)


def uniform_system(n, nrhs, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, n * n), rng.standard_normal(n * nrhs)


def call_dgesv_nrhs(a, b, n, nrhs):
    a, b = a.copy(), b.copy()
    piv = np.zeros(n, dtype=np.uintp)
    st = L.al_dgesv_nrhs(a.ctypes.data_as(_lib._dp), a.size, b.ctypes.data_as(_lib._dp), b.size, n, nrhs,
                         piv.ctypes.data_as(_lib._sp))
    return st, a, b, piv


def test_panel_wait_that_gives_up_is_a_device_error(tiny_wait_limit):
    n = 2048  # eight cooperative panel workgroups exchange every column
    a, b = uniform_system(n, 2, 1)
    st, _, _, _ = call_dgesv_nrhs(a, b, n, 2)
    assert st == ERR_DEVICE, (st, _lib.last_error())
    L.al_debug_set_wait_limit(0)
    st, _, x, _ = call_dgesv_nrhs(a, b, n, 2)
    assert st == 0
    for k in range(2):
        assert O.backward_error(a, x[k * n:(k + 1) * n], b[k * n:(k + 1) * n], n) <= 1e-14


def test_solve_wait_that_gives_up_is_a_device_error():
    n = 1500
    a, b = uniform_system(n, 3, 2)
    st, lu, _, piv = call_dgesv_nrhs(a, b, n, 3)
    assert st == 0
    L.al_debug_set_wait_limit(1)
    try:
        bb = b.copy()
        st = L.al_dgetrs_nrhs(lu.ctypes.data_as(_lib._dp), lu.size, piv.ctypes.data_as(_lib._sp), n,
                              bb.ctypes.data_as(_lib._dp), bb.size, n, 3)
        assert st == ERR_DEVICE, (st, _lib.last_error())
        assert np.isnan(bb).any()  # the blocks that gave up poisoned their rows
    finally:
        L.al_debug_set_wait_limit(0)
    bb = b.copy()
    st = L.al_dgetrs_nrhs(lu.ctypes.data_as(_lib._dp), lu.size, piv.ctypes.data_as(_lib._sp), n,
                          bb.ctypes.data_as(_lib._dp), bb.size, n, 3)
    assert st == 0
    for k in range(3):
        assert O.backward_error(a, bb[k * n:(k + 1) * n], b[k * n:(k + 1) * n], n) <= 1e-14


def test_mixed_solve_wait_that_gives_up_is_a_device_error(tiny_wait_limit):
    n = 1024
    a, b = uniform_system(n, 1, 3)
    bb = b.copy()
    stats = _lib.IrStats()
    st = L.al_dgesv_mixed(a.ctypes.data_as(_lib._dp), a.size, bb.ctypes.data_as(_lib._dp), bb.size, n, 1,
                          _lib.AL_FACTOR_FP32, None, ctypes.byref(stats))
    assert st == ERR_DEVICE, (st, _lib.last_error())


@pytest.mark.parametrize("factor", [serve.FP64, serve.FP32])
def test_server_keeps_no_factor_whose_waits_gave_up(factor):
    n = 1024
    a, b = uniform_system(n, 2, 4)
    with serve.Server(slots=1) as srv:
        L.al_debug_set_wait_limit(1)
        try:
            _, res = srv.wait(srv.submit(a, b, n, 2, serve.GESV, factor, keep=True))
        finally:
            L.al_debug_set_wait_limit(0)
        assert res["status"] == ERR_DEVICE, res
        assert res["factor_id"] == -1
        x, res = srv.wait(srv.submit(a, b, n, 2, serve.GESV, factor, keep=True))
        assert res["status"] == 0 and res["factor_id"] >= 0
        b2 = np.random.default_rng(5).standard_normal(n)
        x2, r2 = srv.wait(srv.submit(None, b2, n, 1, serve.GESV, factor, factor_id=res["factor_id"]))
        assert r2["status"] == 0
        for k in range(2):
            assert O.backward_error(a, x[k * n:(k + 1) * n], b[k * n:(k + 1) * n], n) <= 1e-14
        assert O.backward_error(a, x2, b2, n) <= 1e-14


def test_concurrent_fp64_factorizations_share_the_cus():
    # four FP64 factorizations of order 18000: 71 cooperative panel workgroups each, 284 in all,
    # more than the 128 CUs of half the device; the reservations keep those in flight within it
    n, count = 18000, 4
    systems = [uniform_system(n, 1, 10 + i) for i in range(count)]
    peak = [0]
    stop = threading.Event()

    def watch():
        while not stop.is_set():
            peak[0] = max(peak[0], L.al_debug_coop_units_in_use())
            time.sleep(0.002)

    with serve.Server(hbm_budget_bytes=256 << 30, slots=count) as srv:
        th = threading.Thread(target=watch)
        th.start()
        try:
            tickets = [srv.submit(a, b, n) for a, b in systems]
            results = [srv.wait(t) for t in tickets]
        finally:
            stop.set()
            th.join()
    print(f"peak cooperative units in use: {peak[0]}", flush=True)
    assert 0 < peak[0] <= 128
    for (a, b), (x, res) in zip(systems, results):
        assert res["status"] == 0, res
        assert O.backward_error(a, x, b, n) <= 1e-14
