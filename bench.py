#!/usr/bin/env python3
"""Benchmark: converged grid-cells/s and p50/p99 latency, bf16 served form
against the f32 plain form, on one MI355X.

Workload: one request at a time of the cell_st model (d=128, 4 layers,
4 heads x 32, MLP 512, 27-key stencil = 3x3 periodic in space x 3 in time,
random init N(0, 0.02) from seed 0), on a 64x64 grid with 16 time slices:
5 Gaussian drops (amplitude U[0.05, 0.15], sigma U[2, 5] cells, seeded
positions), slice 0 known, iterated until max|dx| < 1e-4 or 32 iterations.
A step is one whole request: reset the grid to x0 and iterate to the fixed
point. Inputs are resident in HBM before the timed region.

Prints one JSON line. `value` is the bf16 form's converged cells/s;
`fp32_plain` holds the f32 form's figures on the same requests. `roofline`
is measured with HIP events on the launch stream over the timed steps.
`cpu_baseline` times the CPU oracle (a port of the reference forward,
extended to this form) on the box's host cores on a bounded sample.
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

T, NY, NX = 16, 64, 64
CELLS = T * NY * NX
PEAK = {"mfma_f32": 157.3e12, "mfma_bf16": 2.5e15, "hbm": 8.0e12}
CLASSES = ["gemm", "attention", "layernorm", "projection"]


def workload():
    from jacobi2000_amd import checkpoint as ck
    import jacobi2000_amd as j

    st, cj = ck.random_cell_st(dim=128, layers=4, heads=4, ff=512, seed=0, max_iter=32, tol=1e-4)
    grid = ck.drops_grid(T, NY, NX, 5, seed=0)
    known = ck.mask_of("forward", T, NY, NX)
    x0 = j.fill_blank(grid, known)
    meta = ck.meta_of(1.0, 0.09, 9.81, 1.0, 1e-3)
    return st, cj, x0, known, meta


def run_precision(prec, st, cj, x0, known, meta, steps, warmup):
    import torch
    import ctypes
    import jacobi2000_amd as j
    from jacobi2000_amd import _lib

    dev = torch.device("cuda:0")
    model = j.Model.from_bytes(st, cj, precision=prec, device=0)
    d_x0 = torch.from_numpy(x0).to(dev)
    d_grid = torch.empty_like(d_x0)
    d_known = torch.from_numpy(known).to(dev)
    d_meta = torch.from_numpy(meta).to(dev)
    stream = torch.cuda.current_stream().cuda_stream

    def one():
        d_grid.copy_(d_x0)
        return model.iterate_dev(d_grid.data_ptr(), d_known.data_ptr(), d_meta.data_ptr(), T, NY, NX, 32, 1e-4, stream)

    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    timing = (ctypes.c_double * 12)()
    _lib.call("j2k_model_timing", model.handle, 1, None)
    lat, iters = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        it, r, v = one()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
        iters.append(it)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.call("j2k_model_timing", model.handle, 0, timing)
    out = d_grid.cpu().numpy()
    kern = {}
    for c, name in enumerate(CLASSES):
        ms, work, n = timing[3 * c], timing[3 * c + 1], timing[3 * c + 2]
        kern[name] = {"ms_total": ms, "launches": int(n), "work": work}
    model.close()
    lat_ms = np.array(lat) * 1e3
    return {
        "value": CELLS * steps / elapsed,
        "ms_per_step": elapsed * 1e3 / steps,
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "iterations": iters,
        "cell_iterations_per_s": CELLS * sum(iters) / elapsed,
        "resid": float(r),
        "energy_violation": float(v),
        "kernels": kern,
    }, out


def roofline(res, prec):
    k = res["kernels"]
    dom = max(k, key=lambda n: k[n]["ms_total"])
    d = k[dom]
    per_launch_s = d["ms_total"] / max(d["launches"], 1) / 1e3
    work_per_launch = d["work"] / max(d["launches"], 1)
    if dom == "gemm":
        peak = PEAK["mfma_bf16" if prec == "bf16" else "mfma_f32"]
        achieved = work_per_launch / per_launch_s
        return {"kernel": "gemm_" + prec, "bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                "avg_launch_us": per_launch_s * 1e6, "work_per_launch": work_per_launch}
    achieved = work_per_launch / per_launch_s
    return {"kernel": dom + "_" + prec, "bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK["hbm"] / 1e9,
            "unit": "GB/s", "frac": achieved / PEAK["hbm"], "traffic": None,
            "avg_launch_us": per_launch_s * 1e6, "work_per_launch": work_per_launch}


def cpu_baseline(st, cj, x0, known, meta, iters, seconds):
    """The CPU oracle on the same request, all host threads, repeated until
    `seconds` of work (at least one request), with the iteration count the
    GPU f32 run needed."""
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(threads)
    m = O.FormC.load(st, cj)
    done, t0 = 0, time.perf_counter()
    while True:
        m.iterate(x0, known, meta, iters, 1e-4)
        done += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    return {"value": CELLS * done / el, "unit": "converged grid-cells/s", "cores": threads, "kind": "port",
            "sample": f"{done} request(s) of 64x64x16 cells, {iters} iteration(s) each, CPU oracle f32 "
                      f"(restatement of engine::cpu extended to the cell_st form), {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one, and this path has no multi-GPU form "
              "(requests shard across GPUs without communication); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    st, cj, x0, known, meta = workload()
    f32, out32 = run_precision("f32", st, cj, x0, known, meta, args.steps, args.warmup)
    bf16, out16 = run_precision("bf16", st, cj, x0, known, meta, args.steps, args.warmup)
    rel = float(np.linalg.norm(out16 - out32) / np.linalg.norm(out32))
    line = {
        "metric": "converged grid-cells/s + p50/p99 latency, BF16 served vs FP32 plain",
        "value": bf16["value"],
        "unit": "converged grid-cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": bf16["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic: seeded random-init weights N(0,0.02), seeded Gaussian drops",
        "config": {"workload": "cell_st fixed point, one request 64x64 x 16 slices, 27-key stencil, "
                               "model20 shape (d128 L4 H4x32 ff512), slice 0 known, tol 1e-4, cap 32",
                   "cells_per_request": CELLS},
        "p50_ms": bf16["p50_ms"],
        "p99_ms": bf16["p99_ms"],
        "iterations": bf16["iterations"],
        "energy_violation": bf16["energy_violation"],
        "fp32_plain": {k: f32[k] for k in ("value", "ms_per_step", "p50_ms", "p99_ms", "iterations",
                                          "energy_violation")},
        "bf16_vs_fp32_plain": bf16["value"] / f32["value"],
        "bf16_rel_l2_vs_fp32": rel,
        "kernels_bf16": bf16["kernels"],
        "kernels_fp32": f32["kernels"],
        "roofline": roofline(bf16, "bf16"),
        "roofline_fp32": roofline(f32, "f32"),
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(st, cj, x0, known, meta, max(f32["iterations"][0], 1), args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
