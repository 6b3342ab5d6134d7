"""Tuner of the batched solver's launch table (batch_tuning.json beside the library).

For each size bucket of al_dev_batch_solve (0: n <= 64, 1: <= 128, 2: <= 256, 3: <= 512) it
times that bucket's share of the batched workload -- orders log-uniform in [8, 512], half dgesv
(A uniform(-1, 1) + 0.5 n I, 1..8 right-hand sides) and half dposv (G G^T + n I), largest first --
under every candidate (panel width kb, persistent workgroups per CU), in FP64 and FP32 + FP64
refinement, and keeps the configuration with the smallest summed time of the two modes.  The
candidates are what each bucket's LDS allows (al_batch_tune_set rejects the rest).  Writes the
table -- {"table": [{"bucket", "max_n", "kb", "wgs_per_cu"} x 4], ...} first, which the library reads
-- with every measurement beside it.

Run on the GPU:  python algebra_linear_amd/tune_batch.py [--systems 20000] [--out PATH]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

CANDIDATES = {
    0: [(0, w) for w in (1, 2, 3)],
    1: [(16, w) for w in (1, 2, 3, 4, 6)] + [(32, w) for w in (1, 2, 3)],
    2: [(16, w) for w in (1, 2, 3)] + [(32, 1)],
    3: [(16, w) for w in (1, 2)],
}
BUCKET_RANGE = {0: (1, 64), 1: (65, 128), 2: (129, 256), 3: (257, 512)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=20000)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(HERE, "batch_tuning.json"))
    args = ap.parse_args()

    import numpy as np
    import torch

    from algebra_linear_amd import _lib
    lib = _lib.load()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(1988)
    count = args.systems
    ns = np.rint(np.exp(rng.uniform(np.log(8), np.log(512), count))).astype(np.int64)
    kinds = np.where(np.arange(count) % 2 == 0, _lib.AL_BATCH_GESV, _lib.AL_BATCH_POSV)
    nrhs = rng.integers(1, 9, count)
    order = np.argsort(-ns, kind="stable")
    ns, kinds, nrhs = ns[order], kinds[order], nrhs[order]
    table = {"table": [], "device": torch.cuda.get_device_name(0),
             "cus": torch.cuda.get_device_properties(0).multi_processor_count,
             "workload": f"{count} systems, orders log-uniform [8, 512], half dgesv half dposv, 1..8 rhs, largest first",
             "buckets": []}
    for b in range(4):
        lo, hi = BUCKET_RANGE[b]
        sel = (ns >= lo) & (ns <= hi)
        n_, k_, r_ = ns[sel], kinds[sel], nrhs[sel]
        c = len(n_)
        a_off = np.concatenate([[0], np.cumsum(n_ * n_)[:-1]])
        b_off = np.concatenate([[0], np.cumsum(n_ * r_)[:-1]])
        a_len, b_len = int((n_ * n_).sum()), int((n_ * r_).sum())
        g = torch.Generator(device=dev)
        g.manual_seed(1989 + b)
        A0 = torch.rand(a_len, dtype=torch.float64, device=dev, generator=g) * 2 - 1
        B0 = torch.randn(b_len, dtype=torch.float64, device=dev, generator=g)
        for i in range(c):
            n, o = int(n_[i]), int(a_off[i])
            blk = A0[o:o + n * n].view(n, n)
            if k_[i] == _lib.AL_BATCH_POSV:
                blk.copy_(blk @ blk.T + n * torch.eye(n, dtype=torch.float64, device=dev))
            else:
                blk.diagonal().add_(0.5 * n)
        ent = (_lib.BatchEntry * c)(*[_lib.BatchEntry(int(a_off[i]), int(b_off[i]), int(n_[i]), int(r_[i]),
                                                      int(k_[i]), 0) for i in range(c)])
        dE = torch.from_numpy(np.frombuffer(bytes(ent), dtype=np.uint8).copy()).to(dev)
        A, B = torch.empty_like(A0), torch.empty_like(B0)
        info = torch.zeros(c, dtype=torch.int32, device=dev)
        results = []
        for kb, w in CANDIDATES[b]:
            if lib.al_batch_tune_set(b, kb if b else 16, w) != 0:
                continue
            times = {}
            for mode, factor in (("fp64", _lib.AL_FACTOR_FP64), ("fp32_ir", _lib.AL_FACTOR_FP32)):
                best = float("inf")
                for _ in range(args.reps + 1):
                    A.copy_(A0)
                    B.copy_(B0)
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    st = lib.al_dev_batch_solve(c, dE.data_ptr(), A.data_ptr(), a_len, B.data_ptr(), info.data_ptr(),
                                                None, factor, None)
                    torch.cuda.synchronize()
                    if st != 0:
                        raise RuntimeError(f"batch_solve: {st} {_lib.last_error()}")
                    best = min(best, time.perf_counter() - t0)
                if int(info.abs().max().item()) != 0:
                    raise RuntimeError("a system of the tuning workload failed")
                times[mode] = round(best * 1e3, 3)
            results.append({"kb": kb, "wgs_per_cu": w, "ms": times})
            print(f"bucket {b} (n {lo}..{hi}, {c} systems) kb {kb} wgs/CU {w}: {times}", flush=True)
        pick = min(results, key=lambda r: r["ms"]["fp64"] + r["ms"]["fp32_ir"])
        lib.al_batch_tune_set(b, pick["kb"] if b else 16, pick["wgs_per_cu"])
        table["table"].append({"bucket": b, "max_n": hi, "kb": pick["kb"], "wgs_per_cu": pick["wgs_per_cu"]})
        table["buckets"].append({"bucket": b, "max_n": hi, "systems": c, "candidates": results})
        del A0, B0, A, B
        torch.cuda.empty_cache()
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print(json.dumps(table["table"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
