"""Error kinds of the crate's API, as Python exceptions.

* ``LapackError`` mirrors ``enum LapackError { Singular, BadShape, Convergence, .. }``
  (src/lapack/mod.rs:77-96): drivers raise it where the crate returns ``Err(..)``; its
  ``kind`` is one of ``SINGULAR``, ``BAD_SHAPE``, ``CONVERGENCE``.
* ``BlasPanic`` mirrors the crate's BLAS ``assert_eq!`` panics; its message is the
  crate's panic text (e.g. ``"dgemm: a must be m*k"``).
* ``DeviceError`` is a GPU/runtime failure, which the crate (CPU only) cannot have.
"""
from __future__ import annotations

from . import _lib

SINGULAR = "Singular"
BAD_SHAPE = "BadShape"
CONVERGENCE = "Convergence"


class LapackError(Exception):
    def __init__(self, kind: str, message: str = ""):
        super().__init__(f"{kind}: {message}" if message else kind)
        self.kind = kind


class BlasPanic(AssertionError):
    pass


class DeviceError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"device error {code}: {message}")
        self.code = code


def check_lapack(status: int) -> None:
    if status == _lib.AL_OK:
        return
    msg = _lib.last_error()
    if status == _lib.AL_SINGULAR:
        raise LapackError(SINGULAR, msg)
    if status == _lib.AL_BAD_SHAPE:
        raise LapackError(BAD_SHAPE, msg)
    if status == _lib.AL_CONVERGENCE:
        raise LapackError(CONVERGENCE, msg)
    raise DeviceError(status, msg)


def check_blas(status: int) -> None:
    if status == _lib.AL_OK:
        return
    msg = _lib.last_error()
    if status == _lib.AL_BAD_SHAPE:
        raise BlasPanic(msg)
    raise DeviceError(status, msg)
