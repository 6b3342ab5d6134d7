"""Python side of the solve server (``al_server_*``): continuous batching of independent
dgesv / dposv requests on one device, FP64 or FP32 + FP64 refinement, with resident
factorizations for repeated right-hand sides.  Requests take host numpy arrays (column-major
float64, ``a`` n*n and ``b`` n*nrhs); the server writes the solution into a new array.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .errors import DeviceError

GESV, POSV = _lib.AL_BATCH_GESV, _lib.AL_BATCH_POSV
FP64, FP32 = _lib.AL_FACTOR_FP64, _lib.AL_FACTOR_FP32


class Ticket:
    """A submitted request; keeps its host buffers alive until it is waited for."""

    def __init__(self, ticket: int, a, b, x, n: int, nrhs: int):
        self.ticket, self.a, self.b, self.x, self.n, self.nrhs = ticket, a, b, x, n, nrhs


class Server:
    def __init__(self, hbm_budget_bytes: int = 64 << 30, slots: int = 4):
        self._lib = _lib.load()
        self._h = self._lib.al_server_create(int(hbm_budget_bytes), int(slots))
        if not self._h:
            raise DeviceError(_lib.AL_ERR_NO_DEVICE, _lib.last_error())

    def submit(self, a, b, n: int, nrhs: int = 1, kind: int = GESV, factor: int = FP64, keep: bool = False,
               factor_id: int = -1) -> Ticket:
        bb = np.ascontiguousarray(np.asarray(b, dtype=np.float64)).reshape(-1)
        if bb.size != n * nrhs:
            raise ValueError("b must be n*nrhs")
        aa = None
        if factor_id < 0:
            aa = np.ascontiguousarray(np.asarray(a, dtype=np.float64)).reshape(-1)
            if aa.size != n * n:
                raise ValueError("a must be n*n")
        x = np.empty(n * nrhs)
        req = _lib.Request(kind, factor, n, nrhs, aa.ctypes.data if aa is not None else None, bb.ctypes.data,
                           x.ctypes.data, factor_id, 1 if keep else 0, 0)
        t = ctypes.c_uint64()
        st = self._lib.al_server_submit(self._h, ctypes.byref(req), ctypes.byref(t))
        if st != 0:
            raise DeviceError(st, _lib.last_error())
        return Ticket(t.value, aa, bb, x, n, nrhs)

    def wait(self, tk: Ticket):
        """Blocks until the request is done; returns (x, result dict)."""
        r = _lib.Result()
        st = self._lib.al_server_wait(self._h, tk.ticket, ctypes.byref(r))
        if st != 0:
            raise DeviceError(st, _lib.last_error())
        return tk.x, {name: getattr(r, name) for name, _ in _lib.Result._fields_ if name != "reserved"}

    def release(self, factor_id: int) -> None:
        st = self._lib.al_server_release_factor(self._h, int(factor_id))
        if st != 0:
            raise DeviceError(st, _lib.last_error())

    def close(self) -> None:
        if self._h:
            self._lib.al_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
