"""Benchmark of the dense solve path on one MI355X.

A step is one dgesv -- LU with partial pivoting (getrf) plus the triangular solves (getrs) --
of an n x n system with nrhs right-hand sides, on operands already resident in HBM (the
matrix is restored from a pristine device copy before each step; that copy is inside the
timed region).  Throughput is reported as HPL does: (2/3 n^3 + 2 n^2 nrhs) / time.

Prints ONE JSON line.  Extra objects:
  roofline     -- the dominant kernel (the FP64 MFMA GEMM of the trailing updates): its
                  algorithmic flops / its launches' summed duration, from HIP events recorded
                  around every launch on the stream it runs on during the timed steps;
  cpu_baseline -- the CPU oracle (a C restatement of the crate's unblocked dgesv) timed on a
                  bounded sample on one host core;
  accuracy     -- normwise backward error of the last solve, computed on the host in FP64.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 matrix (vendor figure)
PEAK_HBM_GBS = 8000.0


def hpl_flops(n: int, nrhs: int) -> float:
    return 2.0 / 3.0 * n ** 3 + 2.0 * n * n * nrhs


def cpu_baseline(sample_n: int) -> dict:
    import numpy as np

    import oracle as O
    rng = np.random.default_rng(7)
    a = rng.uniform(-1.0, 1.0, sample_n * sample_n)
    b = rng.uniform(-1.0, 1.0, sample_n)
    t0 = time.perf_counter()
    st, _, _ = O.dgesv(a, b)
    dt = time.perf_counter() - t0
    assert st == 0
    return {"value": hpl_flops(sample_n, 1) / dt / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
            "sample": f"oracle dgesv (the crate's unblocked LU + substitution, C restatement) n={sample_n}, "
                      f"nrhs=1, one core, {dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--cpu-sample-n", type=int, default=1536)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one MI355X; --gpus must be 1 (multi-GPU sharding is "
                                   "design-only)"}))
        return 2

    import numpy as np
    import torch

    import algebra_linear_amd as al
    from algebra_linear_amd import _lib

    lib = _lib.load()
    dev = torch.device("cuda:0")
    n, nrhs = args.n, args.nrhs
    g = torch.Generator(device=dev)
    g.manual_seed(1988)
    A0 = (torch.rand((n, n), dtype=torch.float64, device=dev, generator=g) * 2.0 - 1.0)  # column-major view below
    B0 = (torch.rand((nrhs, n), dtype=torch.float64, device=dev, generator=g) * 2.0 - 1.0)
    A = torch.empty_like(A0)
    B = torch.empty_like(B0)
    ipiv = torch.empty(n, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    stream = torch.cuda.Stream(device=dev)  # default priority; everything below is ordered on it
    torch.cuda.set_stream(stream)
    sp = ctypes.c_void_p(stream.cuda_stream)

    def step():
        A.copy_(A0)
        B.copy_(B0)
        st = lib.al_dev_dgetrf(n, n, A.data_ptr(), n, ipiv.data_ptr(), info.data_ptr(), sp)
        if st != 0:
            raise RuntimeError(f"getrf: {st} {_lib.last_error()}")
        st = lib.al_dev_dgetrs(n, nrhs, A.data_ptr(), n, ipiv.data_ptr(), B.data_ptr(), n, sp)
        if st != 0:
            raise RuntimeError(f"getrs: {st} {_lib.last_error()}")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib.al_prof_reset()
    lib.al_prof_enable(1)
    times = []
    t_all0 = time.perf_counter()
    for _ in range(args.steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        step()
        e1.record(stream)
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    torch.cuda.synchronize()
    t_all = time.perf_counter() - t_all0
    lib.al_prof_enable(0)
    if int(info.item()) != 0:
        raise RuntimeError(f"factorization status {int(info.item())}")

    launches, tot_ms, fl, by = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib.al_prof_query(b"dgemm_nn", ctypes.byref(launches), ctypes.byref(tot_ms), ctypes.byref(fl), ctypes.byref(by))
    pl, pms, pfl, pby = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib.al_prof_query(b"panel_getf2", ctypes.byref(pl), ctypes.byref(pms), ctypes.byref(pfl), ctypes.byref(pby))
    lib.al_prof_reset()

    # accuracy of the last solve on the host: A0, x and b in FP64
    a_host = A0.cpu().numpy()            # row-major (n, n) view of the column-major matrix: a_host[j, i] = A[i, j]
    x = B.cpu().numpy()                  # (nrhs, n): column c of X is x[c]
    b = B0.cpu().numpy()
    r = b.T - a_host.T @ x.T
    anorm = np.abs(a_host).sum(axis=0).max()
    be = float(max(np.abs(r[:, c]).max() / (anorm * np.abs(x[c]).max() + np.abs(b[c]).max()) for c in range(nrhs)))

    ms = float(np.mean(times))
    flops = hpl_flops(n, nrhs)
    value = flops / (ms * 1e-3) / 1e12
    avg_ms = tot_ms.value / max(launches.value, 1)
    achieved = fl.value / max(launches.value, 1) / (avg_ms * 1e-3) / 1e12
    out = {
        "metric": "dgesv TFLOP/s (HPL flop count) at FP64 backward error",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: A, B uniform(-1,1) (torch Philox, seed 1988), generated on the device",
        "config": {"workload": f"plain FP64 dgesv n={n} nrhs={nrhs} (getrf + getrs, device-resident)",
                   "n": n, "nrhs": nrhs, "form": "fp64"},
        "solves_per_s": round(1000.0 / ms, 3),
        "p50_ms": round(float(np.percentile(times, 50)), 3),
        "p99_ms": round(float(np.percentile(times, 99)), 3),
        "accuracy": {"backward_error": be, "bound": 1e-14},
        "roofline": {"kernel": "dgemm_mfma<NN> (trailing updates)", "bound": "mfma",
                     "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
                     "launches_per_step": launches.value // max(args.steps, 1),
                     "avg_launch_ms": round(avg_ms, 4),
                     "share_of_step": round(tot_ms.value / (ms * args.steps), 4)},
        "panel": {"launches_per_step": pl.value // max(args.steps, 1),
                  "ms_per_step": round(pms.value / max(args.steps, 1), 3)},
        "wall_s": round(t_all, 3),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample_n)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
