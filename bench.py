"""Benchmark of the dense solve path on one MI355X: mixed-precision dgesv vs plain FP64 dgesv.

Workload (the mixed-precision configuration): A = U diag(sigma) V^T of order n with U, V
Haar-random orthogonal (QR of a Gaussian matrix, signs fixed by R's diagonal), sigma
geometric from 1 down to 1/kappa; B Gaussian, n x nrhs; all generated on the device from a
fixed seed.  A step is one complete solve of A X = B on operands already resident in HBM:

  mixed (headline, --form fp32|bf16): al_dev_dgesv_mixed -- FP32 LU with tournament pivoting
      (trailing updates on FP32 or bf16 matrix cores), FP64 iterative refinement to an FP64
      backward error, FP64 fallback if refinement fails; A is left untouched;
  fp64 (reported beside it): al_dev_dgetrf + al_dev_dgetrs -- the drop-in FP64 path with the
      crate's partial pivoting (A restored from a pristine copy inside the timed region).

Throughput is the HPL / HPL-MxP convention: (2/3 n^3 + 2 n^2 nrhs) FP64-equivalent flops per
second of wall time of a step.  Prints ONE JSON line; extra objects:
  forms        -- every form measured: ms per step, TFLOP/s, refinement steps, fallback,
                  backward error (on the host, FP64) -- and the mixed/fp64 speedup;
  roofline     -- the GEMM doing the O(n^3) work of the headline form (its algorithmic flops
                  / its launches' summed duration, from HIP events around every launch on the
                  stream it runs on, during the timed steps);
  kernels      -- time per step of the instrumented kernel classes;
  cpu_baseline -- the CPU oracle (a C restatement of the crate's unblocked dgesv) timed on a
                  bounded sample on one host core.
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

PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}   # MI355X dense matrix-core peaks
CLASSES = ["sgemm_nn", "bgemm", "dgemm_nn", "tslu_tree", "panel_getf2", "trsv", "residual_gemv"]


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
    ap.add_argument("--kappa", type=float, default=1e2)
    ap.add_argument("--form", choices=["fp32", "bf16"], default="fp32")
    ap.add_argument("--also", default="bf16,fp64", help="other forms measured beside the headline")
    ap.add_argument("--cpu-sample-n", type=int, default=1536)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one MI355X; --gpus must be 1 (multi-GPU sharding is "
                                   "design-only)"}))
        return 2

    import numpy as np
    import torch

    from algebra_linear_amd import _lib, lapack

    lib = _lib.load()
    dev = torch.device("cuda:0")
    n, nrhs = args.n, args.nrhs
    g = torch.Generator(device=dev)
    g.manual_seed(1988)

    def haar():
        q, r = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
        return q * torch.sign(torch.diagonal(r))

    U = haar()
    V = haar()
    sig = args.kappa ** (-torch.arange(n, dtype=torch.float64, device=dev) / max(n - 1, 1))
    # column-major storage of A: a row-major tensor holding A^T
    At = ((U * sig) @ V.T).T.contiguous()
    del U, V
    B0t = torch.randn((nrhs, n), dtype=torch.float64, device=dev, generator=g)   # B^T, row-major
    torch.cuda.empty_cache()
    stream = torch.cuda.Stream(device=dev)  # default priority; everything below is ordered on it
    torch.cuda.set_stream(stream)
    sp = ctypes.c_void_p(stream.cuda_stream)

    def host_berr(x_t: "torch.Tensor") -> float:
        a = At.cpu().numpy()     # a[j, i] = A[i, j]
        x = x_t.cpu().numpy()    # (nrhs, n)
        b = B0t.cpu().numpy()
        r = b.T - a.T @ x.T
        anorm = np.abs(a).sum(axis=0).max()
        return float(max(np.abs(r[:, c]).max() / (anorm * np.abs(x[c]).max() + np.abs(b[c]).max())
                         for c in range(nrhs)))

    def run_form(form: str) -> dict:
        X = torch.zeros_like(B0t)
        stats = _lib.IrStats()
        opt = lapack.ir_options()
        if form == "fp64":
            A = torch.empty_like(At)
            ipiv = torch.empty(n, dtype=torch.int32, device=dev)
            info = torch.zeros(1, dtype=torch.int32, device=dev)

            def step():
                A.copy_(At)
                X.copy_(B0t)
                st = lib.al_dev_dgetrf(n, n, A.data_ptr(), n, ipiv.data_ptr(), info.data_ptr(), sp)
                if st != 0:
                    raise RuntimeError(f"getrf: {st} {_lib.last_error()}")
                st = lib.al_dev_dgetrs(n, nrhs, A.data_ptr(), n, ipiv.data_ptr(), X.data_ptr(), n, sp)
                if st != 0:
                    raise RuntimeError(f"getrs: {st} {_lib.last_error()}")
        else:
            code = lapack.FACTORS[form]

            def step():
                st = lib.al_dev_dgesv_mixed(n, nrhs, At.data_ptr(), n, B0t.data_ptr(), n, X.data_ptr(), n, code,
                                            ctypes.byref(opt), ctypes.byref(stats), sp)
                if st not in (0, 3):
                    raise RuntimeError(f"dgesv_mixed: {st} {_lib.last_error()}")

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        lib.al_prof_reset()
        lib.al_prof_enable(1)
        times = []
        for _ in range(args.steps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            step()
            e1.record(stream)
            e1.synchronize()
            times.append(e0.elapsed_time(e1))
        torch.cuda.synchronize()
        lib.al_prof_enable(0)
        kern = {}
        for c in CLASSES:
            cnt, ms_, fl, by = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            lib.al_prof_query(c.encode(), ctypes.byref(cnt), ctypes.byref(ms_), ctypes.byref(fl), ctypes.byref(by))
            if cnt.value:
                kern[c] = {"launches": cnt.value, "total_ms": ms_.value, "flops": fl.value, "bytes": by.value}
        lib.al_prof_reset()
        if form == "fp64" and int(info.item()) != 0:
            raise RuntimeError(f"factorization status {int(info.item())}")
        ms = float(np.mean(times))
        out = {"ms_per_step": round(ms, 3), "tflops": round(hpl_flops(n, nrhs) / (ms * 1e-3) / 1e12, 3),
               "p50_ms": round(float(np.percentile(times, 50)), 3),
               "p99_ms": round(float(np.percentile(times, 99)), 3),
               "backward_error": host_berr(X), "_kernels": kern, "_times": times}
        if form != "fp64":
            out.update({"refinement_steps": stats.iterations, "fell_back": stats.fell_back,
                        "backward_error_low_precision": stats.backward_error_low})
        return out

    forms = {args.form: run_form(args.form)}
    for f in [f for f in args.also.split(",") if f and f != args.form]:
        forms[f] = run_form(f)

    head = forms[args.form]
    kern = head["_kernels"]
    gemm_class = "bgemm" if args.form == "bf16" else "sgemm_nn"
    gk = kern.get(gemm_class, {"launches": 1, "total_ms": float("nan"), "flops": 0.0})
    peak = PEAK_TFLOPS["bf16" if args.form == "bf16" else "f32"]
    avg_ms = gk["total_ms"] / max(gk["launches"], 1)
    achieved = gk["flops"] / max(gk["launches"], 1) / (avg_ms * 1e-3) / 1e12
    steps = max(args.steps, 1)
    out = {
        "metric": "dgesv TFLOP/s (HPL-MxP FP64-equivalent flops) at FP64 backward error, mixed-IR vs FP64 LU",
        "value": head["tflops"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64" if args.form == "fp32" else "bf16/f32/f64",
        "data": "synthetic: A = U diag(sigma) V^T (Haar U, V from QR of Gaussian; sigma geometric 1..1/kappa), "
                "B Gaussian; torch Philox seed 1988, generated on the device",
        "config": {"workload": f"mixed-precision dgesv ({args.form} factor + FP64 refinement) n={n} nrhs={nrhs} "
                               f"kappa={args.kappa:g}, device-resident", "n": n, "nrhs": nrhs,
                   "kappa": args.kappa, "form": args.form},
        "solves_per_s": round(1000.0 / head["ms_per_step"], 3),
        "p50_ms": head["p50_ms"],
        "p99_ms": head["p99_ms"],
        "accuracy": {"backward_error": head["backward_error"], "bound": 1e-14},
        "forms": {f: {k: v for k, v in d.items() if not k.startswith("_")} for f, d in forms.items()},
        "roofline": {"kernel": f"{gemm_class} (trailing updates of the {args.form} factorization)", "bound": "mfma",
                     "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": None,
                     "launches_per_step": gk["launches"] // steps, "avg_launch_ms": round(avg_ms, 4),
                     "share_of_step": round(gk["total_ms"] / (head["ms_per_step"] * steps), 4)},
        "kernels": {c: {"ms_per_step": round(v["total_ms"] / steps, 3), "launches_per_step": v["launches"] // steps}
                    for c, v in kern.items()},
    }
    if "fp64" in forms:
        out["speedup_vs_fp64"] = round(forms["fp64"]["ms_per_step"] / head["ms_per_step"], 3)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample_n)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
