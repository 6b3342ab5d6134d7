"""GPU test of al_dev_batch_solve_noalloc, the FP64 batched solve that allocates nothing.

It runs the same kernels as al_dev_batch_solve with its scratch (the size buckets' counters and
lists) in a static device array, so on the same inputs its solutions, factors and statuses must
equal al_dev_batch_solve's bit for bit.  Two calls in a row, the second with a smaller batch
(the array keeps the first call's contents, which the second must not see), plus the
oversize rejection.
"""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
al = pytest.importorskip("algebra_linear_amd")
from algebra_linear_amd import _lib  # noqa: E402

LIB = _lib.load()


def make(specs, seed):
    rng = np.random.default_rng(seed)
    a_parts, b_parts, ents = [], [], []
    a_off = b_off = 0
    for n, nrhs, kind in specs:
        if kind == _lib.AL_BATCH_POSV:
            g = rng.uniform(-1, 1, (n, n))
            a = g @ g.T + n * np.eye(n)
        else:
            a = rng.uniform(-1, 1, (n, n)) + 0.5 * n * np.eye(n)
        a_parts.append(a.reshape(-1, order="F"))
        b_parts.append(rng.standard_normal(n * nrhs))
        ents.append(_lib.BatchEntry(a_off, b_off, n, nrhs, kind, 0))
        a_off += n * n
        b_off += n * nrhs
    return np.concatenate(a_parts), np.concatenate(b_parts), ents


def solve(fn_name, A, B, ents, dev, stream):
    dA = torch.from_numpy(A.copy()).to(dev)
    dB = torch.from_numpy(B.copy()).to(dev)
    raw = np.frombuffer(bytes((_lib.BatchEntry * len(ents))(*ents)), dtype=np.uint8).copy()
    dE = torch.from_numpy(raw).to(dev)
    info = torch.full((len(ents),), 99, dtype=torch.int32, device=dev)
    steps = torch.full((len(ents),), 99, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    sp = ctypes.c_void_p(stream.cuda_stream)
    if fn_name == "alloc":
        st = LIB.al_dev_batch_solve(len(ents), dE.data_ptr(), dA.data_ptr(), A.size, dB.data_ptr(), info.data_ptr(),
                                    steps.data_ptr(), _lib.AL_FACTOR_FP64, sp)
    else:
        st = LIB.al_dev_batch_solve_noalloc(len(ents), dE.data_ptr(), dA.data_ptr(), dB.data_ptr(), info.data_ptr(),
                                            steps.data_ptr(), sp)
    assert st == 0, _lib.last_error()
    stream.synchronize()
    return dA.cpu().numpy(), dB.cpu().numpy(), info.cpu().numpy(), steps.cpu().numpy()


def test_noalloc_batch_equals_the_allocating_entry_point_twice():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    first = [(n, 1 + n % 7, kind) for n in (512, 300, 257, 256, 200, 129, 128, 100, 65, 64, 40, 17, 5, 1)
             for kind in (_lib.AL_BATCH_GESV, _lib.AL_BATCH_POSV)]
    second = [(n, 1 + n % 3, kind) for n in (130, 60, 9) for kind in (_lib.AL_BATCH_GESV, _lib.AL_BATCH_POSV)]
    for specs, seed in ((first, 1), (second, 2)):   # the second call is smaller than the first
        A, B, ents = make(specs, seed)
        a1, x1, i1, s1 = solve("alloc", A, B, ents, dev, stream)
        a2, x2, i2, s2 = solve("noalloc", A, B, ents, dev, stream)
        assert list(i1) == [0] * len(ents) and list(s1) == [0] * len(ents)
        np.testing.assert_array_equal(i2, i1)
        np.testing.assert_array_equal(s2, s1)
        np.testing.assert_array_equal(x2, x1)
        np.testing.assert_array_equal(a2, a1)
        for e in ents:   # and the solutions meet the crate's contract
            n = e.n
            a = A[e.a_off:e.a_off + n * n]
            if e.kind == _lib.AL_BATCH_POSV:
                a2d = a.reshape(n, n, order="F")
                a = (np.tril(a2d) + np.tril(a2d, -1).T).reshape(-1, order="F")
            for k in range(e.nrhs):
                o = e.b_off + k * n
                assert O.backward_error(a, x2[o:o + n], B[o:o + n], n) <= 1e-14


def test_noalloc_batch_rejects_more_systems_than_its_scratch_holds():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    st = LIB.al_dev_batch_solve_noalloc(65537, None, None, None, None, None, None)
    assert st == _lib.AL_ERR_ARG
    assert "65536" in _lib.last_error()
