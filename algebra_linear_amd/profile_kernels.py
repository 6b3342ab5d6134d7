"""Profiling driver: runs the library's hot kernels on device-resident data without any PyTorch
kernel, so that rocprofv3 counter collection (--pmc) sees only this library's dispatches.

Device memory comes from the HIP runtime itself (hipMalloc / hipMemcpy through ctypes); inputs
are generated on the host with numpy.  Workloads (--what, comma-separated):
  mixed  -- al_dev_dgesv_mixed, FP32 factor + FP64 refinement, n = 16384, one right-hand side,
            A uniform(-1, 1) (the kernels' launch shapes do not depend on the values);
  mixedfb -- the same with row n-1 replaced by row 0 plus 1e-9 uniform(-1, 1): kappa near 1e9,
            beyond FP32 refinement, so the solve ends in the FP64 LU in the FP32 factor's row
            order;
  dgemm  -- al_dev_dgemm nn, m = n = k = 8192, uniform(-1, 1);
  batch  -- al_dev_batch_solve in FP64 on `--systems` systems of the batched workload's size
            distribution (orders log-uniform in [8, 512], half dgesv, half dposv, 1..8 rhs);
  batch32 -- the same systems in the FP32 + FP64-refinement mode;
  zgemm  -- al_dev_zgemm nn, m = n = k = 4096, entries uniform in the unit square;
  zgetrf -- al_dev_zgetrf, n = 8192, entries uniform in the unit square;
  zpotrf -- al_dev_zpotrf lower, n = 8192, a Hermitian matrix with entries in the unit square
            plus 2n I (positive definite by Gershgorin's theorem);
  dgetrf -- al_dev_dgetrf, n = 8192, uniform(-1, 1), the matrix restored by a copy per call;
  dgetrf_rect -- al_dev_dgetrf on a 16384 x 2048 panel, uniform(-1, 1), restored per call;
  dpotrf -- al_dev_dpotrf, n = 8192, A = G G^T + n I with G uniform(-1, 1), restored per call;
  dgesv64 -- the plain FP64 solve, al_dev_dgetrf + al_dev_dgetrs, n = 16384, one right-hand
            side, uniform(-1, 1), restored per call.
Each runs --reps times after one warm-up.

  python -m algebra_linear_amd.profile_kernels --what mixed,dgemm,batch,zgemm,zgetrf,zpotrf
"""
from __future__ import annotations

import argparse
import ctypes
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from algebra_linear_amd import _lib  # noqa: E402


class Device:
    """hipMalloc'd buffers freed on close; the runtime is the one the library is linked with."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.ptrs = []

    def alloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        if self.hip.hipMalloc(ctypes.byref(p), max(nbytes, 256)) != 0:
            raise MemoryError(f"hipMalloc({nbytes})")
        self.ptrs.append(p.value)
        return p.value

    def upload(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        if self.hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) != 0:  # hipMemcpyHostToDevice
            raise RuntimeError("hipMemcpy")
        return p

    def download(self, p: int, out: np.ndarray) -> np.ndarray:
        if self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) != 0:  # hipMemcpyDeviceToHost
            raise RuntimeError("hipMemcpy")
        return out

    def sync(self):
        self.hip.hipDeviceSynchronize()

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def run(fn, reps: int, dev: Device, label: str) -> None:
    fn()
    dev.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    dev.sync()
    print(f"{label}: {(time.perf_counter() - t0) / reps * 1e3:.2f} ms per call", flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--what", default="mixed,dgemm,batch")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--systems", type=int, default=5000)
    args = ap.parse_args()
    lib = _lib.load()
    dev = Device()
    rng = np.random.default_rng(1988)
    what = set(args.what.split(","))
    for label in ("mixed", "mixedfb"):
        if label not in what:
            continue
        n = 16384
        A = rng.uniform(-1.0, 1.0, n * n)
        if label == "mixedfb":
            Am = A.reshape(n, n)  # Am[j, i] = A[i, j] (column-major)
            Am[:, n - 1] = Am[:, 0] + 1e-9 * rng.uniform(-1.0, 1.0, n)
        dA = dev.upload(A)
        del A
        dB = dev.upload(rng.standard_normal(n))
        dX = dev.alloc(8 * n)
        opt = _lib.IrOptions()
        lib.al_ir_default_options(ctypes.byref(opt))
        stats = _lib.IrStats()

        def mixed():
            st = lib.al_dev_dgesv_mixed(n, 1, dA, n, dB, n, dX, n, _lib.AL_FACTOR_FP32, ctypes.byref(opt),
                                        ctypes.byref(stats), None)
            if st != 0:
                raise RuntimeError(f"dgesv_mixed {st}: {_lib.last_error()}")
        run(mixed, args.reps, dev, f"{label} fp32 n={n}")
        print(f"  refinement steps {stats.iterations}, fell back {stats.fell_back}, backward error "
              f"{stats.backward_error:.2e}", flush=True)
        dev.close()
    if "dgemm" in what:
        n = 8192
        dA = dev.upload(rng.uniform(-1.0, 1.0, n * n))
        dB = dev.upload(rng.uniform(-1.0, 1.0, n * n))
        dC = dev.alloc(8 * n * n)

        def gemm():
            if lib.al_dev_dgemm(0, 0, n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n, None) != 0:
                raise RuntimeError(_lib.last_error())
        run(gemm, args.reps, dev, f"dgemm nn {n}")
        dev.close()
    if "batch" in what or "batch32" in what:
        count = args.systems
        ns = np.rint(np.exp(rng.uniform(np.log(8), np.log(512), count))).astype(np.int64)
        kinds = np.where(np.arange(count) % 2 == 0, _lib.AL_BATCH_GESV, _lib.AL_BATCH_POSV)
        nrhs = rng.integers(1, 9, count)
        order = np.argsort(-ns, kind="stable")
        ns, kinds, nrhs = ns[order], kinds[order], nrhs[order]
        a_off = np.concatenate([[0], np.cumsum(ns * ns)[:-1]])
        b_off = np.concatenate([[0], np.cumsum(ns * nrhs)[:-1]])
        A = np.empty(int((ns * ns).sum()))
        for i in range(count):
            n = int(ns[i])
            g = rng.uniform(-1.0, 1.0, (n, n))
            blk = g @ g.T + n * np.eye(n) if kinds[i] == _lib.AL_BATCH_POSV else g + 0.5 * n * np.eye(n)
            A[a_off[i]:a_off[i] + n * n] = blk.reshape(-1, order="F")
        B = rng.standard_normal(int((ns * nrhs).sum()))
        ent = (_lib.BatchEntry * count)(*[_lib.BatchEntry(int(a_off[i]), int(b_off[i]), int(ns[i]), int(nrhs[i]),
                                                          int(kinds[i]), 0) for i in range(count)])
        dE = dev.upload(np.frombuffer(bytes(ent), dtype=np.uint8))
        dA0 = dev.upload(A)
        dA = dev.alloc(A.nbytes)
        dB = dev.upload(B)
        dInfo = dev.alloc(4 * count)

        dB0 = dev.upload(B)
        for mode, factor in (("batch", _lib.AL_FACTOR_FP64), ("batch32", _lib.AL_FACTOR_FP32)):
            if mode not in what:
                continue

            def batch():  # the FP64 mode factors in place: each call starts from the pristine matrices
                if dev.hip.hipMemcpy(dA, dA0, A.nbytes, 3) != 0 or dev.hip.hipMemcpy(dB, dB0, B.nbytes, 3) != 0:
                    raise RuntimeError("hipMemcpy")
                if lib.al_dev_batch_solve(count, dE, dA, A.size, dB, dInfo, None, factor, None) != 0:
                    raise RuntimeError(_lib.last_error())
            run(batch, args.reps, dev, f"batched {'fp64' if mode == 'batch' else 'fp32+ir'}, {count} systems")
            info = dev.download(dInfo, np.zeros(count, dtype=np.int32))
            if np.any(info != 0):
                raise RuntimeError("a batched system failed")
        dev.close()
    for name, (m, n) in (("dgetrf", (8192, 8192)), ("dgetrf_rect", (16384, 2048)), ("dgesv64", (16384, 16384))):
        if name not in what:
            continue
        A = rng.uniform(-1.0, 1.0, m * n)
        dA0, dA = dev.upload(A), dev.alloc(A.nbytes)
        dPiv, dInfo = dev.alloc(4 * n), dev.alloc(4)
        dB0 = dev.upload(rng.standard_normal(m))
        dB = dev.alloc(8 * m)
        del A

        def lu(m=m, n=n, dA0=dA0, dA=dA, dPiv=dPiv, dInfo=dInfo, dB0=dB0, dB=dB, solve=name == "dgesv64"):
            if dev.hip.hipMemcpy(dA, dA0, 8 * m * n, 3) != 0:
                raise RuntimeError("hipMemcpy")
            if lib.al_dev_dgetrf(m, n, dA, m, dPiv, dInfo, None) != 0:
                raise RuntimeError(_lib.last_error())
            if solve:
                if dev.hip.hipMemcpy(dB, dB0, 8 * m, 3) != 0:
                    raise RuntimeError("hipMemcpy")
                if lib.al_dev_dgetrs(n, 1, dA, m, dPiv, dB, n, None) != 0:
                    raise RuntimeError(_lib.last_error())
        run(lu, args.reps, dev, f"{name} {m}x{n}")
        if dev.download(dInfo, np.zeros(1, dtype=np.int32))[0] != 0:
            raise RuntimeError(f"{name} reported a zero pivot")
        dev.close()
    if "dpotrf" in what:
        n = 8192
        G = rng.uniform(-1.0, 1.0, (n, n))
        A = (G @ G.T + n * np.eye(n)).reshape(-1)
        del G
        dA0, dA = dev.upload(A), dev.alloc(A.nbytes)
        dInfo = dev.alloc(4)
        del A

        def potrf():
            if dev.hip.hipMemcpy(dA, dA0, 8 * n * n, 3) != 0:
                raise RuntimeError("hipMemcpy")
            if lib.al_dev_dpotrf(n, dA, n, dInfo, None) != 0:
                raise RuntimeError(_lib.last_error())
        run(potrf, args.reps, dev, f"dpotrf {n}")
        if dev.download(dInfo, np.zeros(1, dtype=np.int32))[0] != 0:
            raise RuntimeError("dpotrf reported a non-positive pivot")
        dev.close()
    one, zero = _lib.Complex.of(1.0), _lib.Complex.of(0.0)

    def cunit(count):
        return rng.uniform(-1.0, 1.0, count) + 1j * rng.uniform(-1.0, 1.0, count)
    if "zgemm" in what:
        n = 4096
        dA, dB = dev.upload(cunit(n * n)), dev.upload(cunit(n * n))
        dC = dev.alloc(16 * n * n)

        def zgemm():
            if lib.al_dev_zgemm(0, 0, n, n, n, one, dA, n, dB, n, zero, dC, n, None) != 0:
                raise RuntimeError(_lib.last_error())
        run(zgemm, args.reps, dev, f"zgemm nn {n}")
        dev.close()
    if "zgetrf" in what or "zpotrf" in what:
        n = 8192
        dInfo = dev.alloc(4)
        dPiv = dev.alloc(4 * n)
        if "zgetrf" in what:
            A = cunit(n * n)
            dA0, dA = dev.upload(A), dev.alloc(A.nbytes)

            def zgetrf():
                if dev.hip.hipMemcpy(dA, dA0, A.nbytes, 3) != 0:
                    raise RuntimeError("hipMemcpy")
                if lib.al_dev_zgetrf(n, n, dA, n, dPiv, dInfo, None) != 0:
                    raise RuntimeError(_lib.last_error())
            run(zgetrf, args.reps, dev, f"zgetrf {n}")
            if dev.download(dInfo, np.zeros(1, dtype=np.int32))[0] != 0:
                raise RuntimeError("zgetrf reported a zero pivot")
            del A
        if "zpotrf" in what:
            H = cunit(n * n).reshape(n, n)
            H = 0.5 * (H + H.conj().T) + 2.0 * n * np.eye(n)
            A = np.ascontiguousarray(H.T).reshape(-1)
            del H
            dA0, dA = dev.upload(A), dev.alloc(A.nbytes)

            def zpotrf():
                if dev.hip.hipMemcpy(dA, dA0, A.nbytes, 3) != 0:
                    raise RuntimeError("hipMemcpy")
                if lib.al_dev_zpotrf(0, n, dA, n, dInfo, None) != 0:
                    raise RuntimeError(_lib.last_error())
            run(zpotrf, args.reps, dev, f"zpotrf lower {n}")
            if dev.download(dInfo, np.zeros(1, dtype=np.int32))[0] != 0:
                raise RuntimeError("zpotrf reported a non-positive pivot")
            del A
        dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
