#!/usr/bin/env python3
"""Benchmark of the plain GPU path on one MI355X: Monte Carlo model build
(move + LDS radix sort/count-by-key + CSR) followed by fused SpMV value iteration to
residual 1e-6, on the 100x100 x nt=120, 16-heading, 5-mode, 2500-realization workload.

One step = one end-to-end build + solve with the fields already resident in HBM.
Prints one JSON line (rank 0). `value` = Monte Carlo transitions per second of the
whole step (lanes / step time). Per-kernel times come from HIP events recorded inside
the library on the stream the kernels run on.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBPS = 8000.0
PEAK_FP32_TFLOPS = 157.3


def _cpu_baseline(w, fields, budget_s: float):
    """The oracle's OpenMP build on the host cores over the first slices of the same
    workload (bounded sample)."""
    import oracle
    threads = min(16, os.cpu_count() or 1)
    p = w.params
    lanes_per_slice = p.nc * p.na * p.nrzns
    t0 = time.perf_counter()
    oracle.build(p, fields, 0, 1, threads)
    t1 = time.perf_counter() - t0
    n = max(1, min(p.nt - 2, int(budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    oracle.build(p, fields, 1, 1 + n, threads)
    dt = time.perf_counter() - t0
    return {
        "value": lanes_per_slice * n / dt, "unit": "transitions/s", "cores": threads,
        "kind": "port",
        "sample": f"oracle build (OpenMP, {threads} threads) of time slices 1..{n} of the same "
                  f"workload: {lanes_per_slice * n:.3e} lanes in {dt:.1f} s",
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="C2")
    ap.add_argument("--thresh", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--skip-vi", action="store_true", help="profiling aid: build only")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              "(run with --gpus 1)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import e2e_gpu_dp_amd as e2e
    from e2e_gpu_dp_amd.configs import workload

    if e2e.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    w = workload(args.workload)
    p = w.params
    print(f"[bench] generating fields for {w.name}", file=sys.stderr, flush=True)
    fields = w.fields()
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in fields.items()}
    torch.cuda.synchronize()

    def step():
        m = e2e.build_model(p, dev)
        st = None
        if not args.skip_vi:
            _, _, st = m.solve(thresh=args.thresh, sweeps_per_check=1, max_sweeps=p.nt + 2)
        info = m.info
        m.close()
        return info, st

    for i in range(args.warmup):
        step()
        print(f"[bench] warmup {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    infos = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        infos.append(step())
        print(f"[bench] step {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    lanes = infos[0][0]["lanes"]

    move_ms = float(np.mean([i["move_kernel_ms"] / max(i["move_launches"], 1) for i, _ in infos]))
    flops_per_lane = 4 * p.nmodes + 18
    build_tflops = lanes * flops_per_lane / (move_ms * 1e-3) / 1e12
    build_ms = float(np.mean([i["build_ms"] for i, _ in infos]))
    out = {
        "metric": "MC transitions/s, VI GB/s, batched solves/s + p99 latency vs plain at 1e-5",
        "value": lanes * args.steps / elapsed,
        "unit": "transitions/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{w.name}: plain build+SpVI, {p.gsize}x{p.gsize} grid x nt={p.nt}, "
                               f"{p.na} headings, {p.nmodes} DO modes, {p.nrzns} realizations",
                   "gsize": p.gsize, "nt": p.nt, "na": p.na, "nrzns": p.nrzns, "nmodes": p.nmodes,
                   "states": p.ns, "nnz": infos[0][0]["nnz"], "lanes": lanes},
        "build": {"transitions_per_s": lanes / (build_ms * 1e-3), "build_ms": build_ms,
                  "move_kernel_ms": move_ms, "scan_ms": infos[0][0]["scan_ms"],
                  "csr_ms": infos[0][0]["csr_ms"]},
        "roofline": {"kernel": "mc_move_count", "bound": "valu", "achieved": build_tflops,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": build_tflops / PEAK_FP32_TFLOPS, "traffic": None,
                     "flops_per_lane": flops_per_lane},
    }
    if not args.skip_vi:
        st = infos[-1][1]
        per = float(np.mean([s["sweep_kernel_ms_per_sweep"] for _, s in infos]))
        gbps = st["bytes_per_sweep"] / (per * 1e-3) / 1e9
        out["vi"] = {"sweeps": st["sweeps"], "solve_ms": float(np.mean([s["solve_ms"] for _, s in infos])),
                     "sweep_kernel_ms": per, "bytes_per_sweep": st["bytes_per_sweep"],
                     "gbps": gbps, "lanes_per_row": st["lanes_per_row"]}
        out["roofline_vi"] = {"kernel": "spmv_vi_sweep", "bound": "hbm", "achieved": gbps,
                              "peak": PEAK_HBM_GBPS, "unit": "GB/s", "frac": gbps / PEAK_HBM_GBPS,
                              "traffic": None}
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = _cpu_baseline(w, fields, args.cpu_budget)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
