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
CLASSES = ["gemm", "attention", "layernorm", "projection", "rows", "rows_tail", "rows_embed"]


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
    # the copy and the library's launches share one stream of their own
    side = torch.cuda.Stream()
    stream = side.cuda_stream

    def reset():  # the call writes its result into the grid: each request starts from x0
        with torch.cuda.stream(side):
            d_grid.copy_(d_x0)

    def run():
        return model.iterate_dev(d_grid.data_ptr(), d_known.data_ptr(), d_meta.data_ptr(), T, NY, NX, 32, 1e-4, stream)

    def one():
        reset()
        return run()

    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    # value: the production path (per-kernel timing off, so the iterations
    # replay from the graphs captured in the warmup); each request's clock
    # starts with its input grid in place in HBM and stops when its stream
    # is done
    lat, iters = [], []
    for _ in range(steps):
        reset()
        side.synchronize()
        ts = time.perf_counter()
        it, r, v = run()
        side.synchronize()
        lat.append(time.perf_counter() - ts)
        iters.append(it)
    torch.cuda.synchronize()
    elapsed = sum(lat)
    out = d_grid.cpu().numpy()
    # kernel classes: a separate pass with per-kernel events (eager launches)
    timing = (ctypes.c_double * (3 * len(CLASSES)))()
    _lib.call("j2k_model_timing", model.handle, 1, None)
    t1 = time.perf_counter()
    timed_steps = max(1, min(steps, 10))
    for _ in range(timed_steps):
        one()
        side.synchronize()
    eager_elapsed = time.perf_counter() - t1
    _lib.call("j2k_model_timing", model.handle, 0, timing)
    kern = {}
    for c, name in enumerate(CLASSES):
        ms, work, n = timing[3 * c], timing[3 * c + 1], timing[3 * c + 2]
        kern[name] = {"ms_total": ms, "launches": int(n), "work": work}
    tuning = model.autotune(0) if prec != "f32" else None
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
        "eager_timed_ms_per_step": eager_elapsed * 1e3 / timed_steps,
        "autotune": tuning,
    }, out


def run_iterating(meta, steps, warmup, headline, no_error):
    """The single-request workload with an input-sensitive checkpoint of the
    same shape (N(0, 0.1) from seed 5, energy bound 1.05), on moving drops
    (the slices differ): the fixed point takes several iterations, so the
    stall rule, graph replay across iterations and convergence are timed.
    Per form: ms per request, iterations, converged cells/s; error against
    the f64 oracle on exact weights and on the form's bf16 weights."""
    from jacobi2000_amd import checkpoint as ck
    import jacobi2000_amd as j

    st, cj = ck.random_cell_st(seed=5, std=0.1, r_max=1.05)
    g = ck.moving_drops_grid(T, NY, NX, 5, 11)
    known = ck.mask_of("forward", T, NY, NX)
    x0 = j.fill_blank(g, known)
    out = {"workload": "64x64 x 16, model20 shape with N(0,0.1) weights (seed 5), energy_r_max 1.05, five moving "
                       "drops, slice 0 known, tol 1e-4, cap 32"}
    res = {}
    for prec in ("f32", "bf16", "bf16x2"):
        r, grid = run_precision(prec, st, cj, x0, known, meta, steps, warmup)
        res[prec] = (r, grid)
        out[prec] = {k: r[k] for k in ("value", "ms_per_step", "p50_ms", "p99_ms", "resid", "energy_violation")}
        out[prec]["iterations"] = r["iterations"][-1]
    out["served_vs_plain"] = res[headline][0]["value"] / res["f32"][0]["value"]
    if not no_error:
        from oracle import oracle as O

        O.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        cpu = O.FormC.load(st, cj)
        wq = {}
        for name, a in cpu.w.items():
            parts = name.split(".")
            q = len(parts) == 4 and parts[2] in ("q", "k", "v", "proj", "ff1", "ff2") and parts[3] == "weight"
            wq[name] = ((ck.f32_to_bf16_bits(a).astype(np.uint32) << 16).view(np.float32) if q else a)
        cpu_bf = O.FormC(cpu.cfg, wq)
        cache = {}
        for prec, (r, grid) in res.items():
            its = r["iterations"][-1]
            x0d = x0.astype(np.float64)
            for label, oracle in (("exact_weights", cpu), ("bf16_weights", cpu_bf)):
                if label == "bf16_weights" and prec == "f32":
                    continue
                if (label, its) not in cache:
                    cache[(label, its)] = oracle.iterate(x0, known, meta, its, 0.0, f64=True)[0]
                ref = cache[(label, its)]
                d = np.linalg.norm((grid - ref).reshape(3, -1), axis=1)
                inc = np.linalg.norm((ref - x0d).reshape(3, -1), axis=1)
                out[prec][f"error_vs_f64_oracle_{label}"] = {
                    "rel_l2": float(np.linalg.norm(grid - ref) / np.linalg.norm(ref)),
                    "rel_to_increment_per_field": [float(a / max(b, 1e-30)) for a, b in zip(d, inc)]}
    return out


def note(msg):
    """Progress line on stderr (the JSON line alone goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


# launch shape of the row-panel kernel (the autotuner's rows_fused pick) ->
# the shape's template arguments: waves along tokens, along columns, MLP chunk
ROWS_SHAPES = {1: "1, 4, 128", 2: "1, 8, 128", 3: "1, 4, 64", 4: "2, 2, 128"}


def kernel_symbol(cls, prec, rows_pick=1):
    """Name (a substring of it) of the kernel a timing class measures, as the
    profiler shows it; rows_pick is the launch shape the autotuner chose."""
    split = "true" if prec == "bf16x2" else "false"
    shape = ROWS_SHAPES.get(rows_pick, ROWS_SHAPES[1])
    return {"gemm": "j2k::k_gemm_f32" if prec == "f32" else "k_rows_embed_qkv",
            "attention": "j2k::k_attn_cell", "layernorm": "j2k::k_layernorm", "projection": "j2k::k_cell_project",
            "rows": f"k_rows<{split}, true, true, false, {shape}, false, false>",
            "rows_tail": f"k_rows<{split}, true, false, false, 1, 4, 128, true, false>",
            "rows_embed": f"k_rows<{split}, true, true, false, 1, 4, 128, false, true>"}[cls]


def rows_pick_of(res):
    """The autotuner's row-panel launch shape at the largest tuned bucket."""
    t = res.get("autotune") or {}
    b = t.get("buckets") or []
    return b[-1].get("rows_fused", 1) if b else 1


def roofline(res, prec, traffic=None):
    """Roofline of the dominant kernel class (most event-timed time): GEMMs
    against the MFMA peak of their precision, the rest (attention, the
    row-panel kernel, norms, projection) against HBM with algorithmic bytes."""
    k = res["kernels"]
    dom = max(k, key=lambda n: k[n]["ms_total"])
    d = k[dom]
    per_launch_s = d["ms_total"] / max(d["launches"], 1) / 1e3
    work_per_launch = d["work"] / max(d["launches"], 1)
    achieved = work_per_launch / per_launch_s
    out = {"kernel": kernel_symbol(dom, prec, rows_pick_of(res)), "class": dom, "avg_launch_us": per_launch_s * 1e6,
           "launches": d["launches"], "work_per_launch": work_per_launch}
    if dom == "gemm":
        peak = PEAK["mfma_f32" if prec == "f32" else "mfma_bf16"]
        out.update(bound="mfma", achieved=achieved / 1e12, peak=peak / 1e12, unit="TFLOP/s", frac=achieved / peak)
    else:
        out.update(bound="hbm", achieved=achieved / 1e9, peak=PEAK["hbm"] / 1e9, unit="GB/s",
                   frac=achieved / PEAK["hbm"])
    out["traffic"] = traffic
    return out


def measure_traffic(args, symbol, force=None):
    """HBM bytes per launch of `symbol` at the workload's size: two
    rocprofv3 --pmc runs of this script (FETCH_SIZE, then WRITE_SIZE, KiB per
    dispatch) on the headline form, each scaled by what the same counter reads
    for a copy of a known 256 MiB buffer in the same run. Returns
    (bytes, detail) or (None, reason)."""
    import csv
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if not exe:
        return None, "rocprofv3 not on PATH"
    child = [sys.executable, str(ROOT / "bench.py"), "--forms", args.headline, "--headline", args.headline,
             "--steps", "1", "--warmup", "1", "--no-cpu-baseline", "--no-oracle-error", "--no-traffic",
             "--stream", "0", "--rollouts", "0", "--wide", "0", "--copy-probe", "256", "--no-plain"]
    probe_kib = 256 * 1024.0
    got = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        with tempfile.TemporaryDirectory(dir="/tmp") as td:
            cmd = [exe, "--pmc", ctr, "--output-format", "csv", "-d", td, "-o", "run", "--"] + child
            try:
                env = dict(os.environ, TMPDIR="/tmp")
                if force:  # the launch shape the measured run chose
                    env["J2K_TUNE_FORCE"] = force
                r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, timeout=240)
            except subprocess.TimeoutExpired:
                return None, f"rocprofv3 --pmc {ctr} timed out"
            files = sorted(Path(td).rglob("*counter_collection.csv"))
            if not files:
                return None, f"rocprofv3 --pmc {ctr} wrote no counters (exit {r.returncode})"
            rows = [row for f in files for row in csv.DictReader(open(f)) if row["Counter_Name"] == ctr]
        mine = [row for row in rows if symbol in row["Kernel_Name"]]
        copies = [row for row in rows if "elementwise" in row["Kernel_Name"] or "copy" in row["Kernel_Name"].lower()]
        if not mine or not copies:
            return None, f"no {symbol} or copy dispatch in the {ctr} run"
        gmax = max(int(row["Grid_Size"]) for row in mine)
        vals = [float(row["Counter_Value"]) for row in mine if int(row["Grid_Size"]) == gmax]
        copy_kib = max(float(row["Counter_Value"]) for row in copies)
        got[ctr] = (sum(vals) / len(vals), copy_kib / probe_kib, len(vals), gmax)
    fetch, fscale, n, grid = got["FETCH_SIZE"]
    write, wscale, _, _ = got["WRITE_SIZE"]
    total = (fetch / fscale + write / wscale) * 1024.0
    return total, {"fetch_kib": fetch, "write_kib": write, "fetch_scale": fscale, "write_scale": wscale,
                   "dispatches": n, "grid_size": grid, "probe": "256 MiB torch copy"}


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


# The reference's published figures for its own forward (form P, model20):
# crates/engine GpuModel::step on an Apple M2 (docs/runs/2026-10-12-engine-
# webgpu-perf.md:29-39) and the rollout CLI on a Linux desktop with an RTX
# 3080 (docs/runs/2026-10-07-inference-lean.md:157-160).
REF_FORM_P_MS = {"webgpu_m2_step": 7.97, "webgpu_vulkan_rollout_step": 17.40, "cpu_rollout_step": 108.87}
FORM_P_GFLOP = 0.578  # per forward at model20 shape, 64x64 grid, patch 4


def _timed(fn, seconds, at_least=1):
    done, t0 = 0, time.perf_counter()
    while done < at_least or time.perf_counter() - t0 < seconds:
        fn()
        done += 1
    return (time.perf_counter() - t0) / done, done


def run_form_p(steps, warmup, cpu_seconds):
    """The reference's own forward (form P) at model20 shape: 64x64 grid,
    patch 4 (256 tokens), d 128, 4 layers, 4 heads x 32, MLP 512, 3x3 token
    stencil, random init N(0, 0.02) seed 0; the rollout CLI's `drop` case
    (main.rs:63-78). Times j2k_step with its host copies (state in, next
    state out: the reference's GpuModel::step also uploads and reads back)
    and the device rollout loop with the mass projection (main.rs:117-121),
    per step; the CPU oracle's forward on all host threads and on one."""
    import ctypes

    import torch
    import jacobi2000_amd as j
    from jacobi2000_amd import _lib
    from jacobi2000_amd import checkpoint as ck

    st, cj = ck.random_patch2d(seed=0)
    n = 64 * 64
    h = np.full((64, 64), 1.0, np.float32)
    ck.add_drop(h, 32.0, 32.0, 0.05, 4.0)
    x = np.concatenate([(h - np.float32(1.0)).reshape(-1), np.zeros(2 * n, np.float32)])
    meta = ck.meta_of(1.0, 0.09, 9.81, 1.0, 1e-3)
    out = {"workload": "reference forward (patch2d) at model20 shape: 64x64 grid, patch 4, d128 L4 H4x32 ff512, "
                       "3x3 token stencil, drop case",
           "reference_ms_per_step": REF_FORM_P_MS, "gflop_per_forward": FORM_P_GFLOP}
    dev = torch.device("cuda:0")
    for prec in ("f32", "bf16x2"):
        model = j.Model.from_bytes(st, cj, precision=prec)
        for _ in range(warmup):
            model.step(x, meta)
        step_s, _ = _timed(lambda: model.step(x, meta), 0.0, at_least=max(steps, 10))
        d_x = torch.from_numpy(x).to(dev)
        d_meta = torch.from_numpy(meta).to(dev)
        loop = 50
        torch.cuda.synchronize()

        def rollout():
            _lib.call("j2k_rollout_dev", model.handle, ctypes.c_void_p(d_x.data_ptr()),
                      ctypes.c_void_p(d_meta.data_ptr()), loop, 1, 0, None, None)
            torch.cuda.synchronize()

        rollout()
        roll_s, reps = _timed(rollout, 0.0, at_least=3)
        model.close()
        out[prec] = {"step_ms": step_s * 1e3, "rollout_ms_per_step": roll_s * 1e3 / loop,
                     "rollout_steps": loop * reps, "gflops_per_s": FORM_P_GFLOP / (roll_s / loop),
                     "vs_webgpu_m2": REF_FORM_P_MS["webgpu_m2_step"] / (step_s * 1e3),
                     "vs_vulkan_rollout": REF_FORM_P_MS["webgpu_vulkan_rollout_step"] / (roll_s * 1e3 / loop)}
    if cpu_seconds > 0:
        from oracle import oracle as O

        cpu = O.FormP.load(st, cj)
        legs = []
        for threads in (int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), 1):
            O.set_threads(threads)
            t, done = _timed(lambda: cpu.forward(x, meta), cpu_seconds, at_least=2)
            legs.append({"cores": threads, "forward_ms": t * 1e3, "forwards": done})
        O.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        out["cpu_oracle"] = {"kind": "port", "what": "CPU restatement of engine::cpu (cpu.rs:249-352), f32",
                             "legs": legs}
    return out


def cpu_legs(st, cj, meta, x0, known):
    """The CPU oracle (f32 restatement of engine::cpu extended to the cell_st
    form) on one iteration of the 32x32 x 8 request (centred drop) and of the
    64x64 x 16 request, on all host threads and on one: ms per iteration and
    cell-iterations/s."""
    from jacobi2000_amd import checkpoint as ck
    from oracle import oracle as O

    cpu = O.FormC.load(st, cj)
    g1 = ck.drops_grid(8, 32, 32, 1, 0, centred=True)
    k1 = ck.mask_of("forward", 8, 32, 32)
    from oracle.oracle import fill_blank

    reqs = {"32x32x8": (fill_blank(g1, k1), k1), "64x64x16": (x0, known)}
    all_threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {"kind": "port", "what": "CPU restatement of engine::cpu extended to the cell_st form, f32, one iteration"}
    for name, (x, k) in reqs.items():
        cells = x[0].size
        legs = []
        for threads in (all_threads, 1):
            O.set_threads(threads)
            t, done = _timed(lambda: cpu.iterate(x, k, meta, 1, 0.0), 1.0 if threads > 1 else 0.0, at_least=1)
            legs.append({"cores": threads, "ms_per_iteration": t * 1e3, "cell_iterations_per_s": cells / t,
                         "iterations_timed": done})
        out[name] = legs
    O.set_threads(all_threads)
    return out


def run_solver(dataset_states, seconds_cpu):
    """The shallow-water solver on the device (crates/solvers/src/
    shallow_water.rs), bit-identical to the reference's f32 step, in three
    shapes, each timed with HIP events on the launch stream:
      - dataset: dump-dataset's loop (dump_dataset.rs:168-193), B states of
        64x64 (nu 1e-3), 80 ticks each, every frame written (resident
        kernel, one launch); HBM carries 12 B per cell per frame plus the
        state in and out;
      - truth partner: one 64x64 state, 50 ticks (the rollout CLI's solver);
      - large grid: one 4096x4096 state, 20 steps, one HBM launch per step (row march)
        moving 24 B per cell (h, u, v in and out).
    The CPU leg runs the oracle's restatement of ShallowWater::step on one
    core at 64x64."""
    import torch

    from jacobi2000_amd import solvers as sw
    from oracle import oracle as O

    dev = torch.device("cuda:0")
    # a stream of its own: torch's default stream is handle 0, which the
    # C ABI reads as "the solver's own stream"
    stream = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(stream):
        return _run_solver(dataset_states, seconds_cpu, sw, O, dev, stream)


def _run_solver(dataset_states, seconds_cpu, sw, O, dev, stream):
    import ctypes

    import torch

    from jacobi2000_amd._lib import call

    out = {}

    def timed(fn, reps=1):
        fn()  # warmup
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / 1e3 / reps

    def state(p, batch, seed):
        rng = np.random.default_rng(seed)
        h = np.full((batch, p.cells()), p.depth, np.float32)
        for i in range(batch):
            h[i] += sw.drop_field(p, rng.uniform(8, p.nx - 8), rng.uniform(8, p.ny - 8), rng.uniform(0.02, 0.08),
                                  rng.uniform(2.5, 6.0)).reshape(-1)
        z = np.zeros_like(h)
        return [torch.from_numpy(a).to(dev) for a in (h, z, z.copy())]

    # dataset
    p = sw.Params(nx=64, ny=64, nu=1e-3)
    B, ticks, n = dataset_states, 80, 64 * 64
    s = sw.ShallowWater(p)
    h0, u0, v0 = state(p, B, 20)
    h, u, v = h0.clone(), u0.clone(), v0.clone()
    frames = ticks + 1
    traj = torch.empty(B * 3 * frames * n, dtype=torch.float32, device=dev)
    tk = torch.full((B,), ticks, dtype=torch.int32, device=dev)
    off = torch.arange(B, dtype=torch.int64, device=dev) * (3 * frames * n)

    def dataset():
        h.copy_(h0), u.copy_(u0), v.copy_(v0)
        call("j2k_sw_trajectories_dev", s._h, ctypes.c_void_p(h.data_ptr()), ctypes.c_void_p(u.data_ptr()),
             ctypes.c_void_p(v.data_ptr()), B, ctypes.c_void_p(tk.data_ptr()), ticks, ctypes.c_void_p(off.data_ptr()),
             ctypes.c_void_p(traj.data_ptr()), None, None, None, ctypes.c_void_p(stream.cuda_stream))

    copy_s = timed(lambda: (h.copy_(h0), u.copy_(u0), v.copy_(v0)), 3)
    t = timed(dataset, 3) - copy_s
    bytes_ = B * n * 4 * (3 * frames + 6)
    out["dataset"] = {"workload": f"{B} states of 64x64, nu 1e-3, {ticks} ticks, all {frames} frames written",
                      "kernel": "k_sw_resident<8, in place>", "ms": t * 1e3,
                      "cell_ticks_per_s": B * n * ticks / t, "frames_per_s": B * ticks / t,
                      "hbm_bytes": bytes_, "achieved_GBps": bytes_ / t / 1e9, "frac_hbm": bytes_ / t / PEAK["hbm"]}
    del traj
    # truth partner: one state, 50 ticks, one launch
    h1, u1, v1 = state(p, 1, 1)
    t = timed(lambda: s.steps_dev(h1.data_ptr(), u1.data_ptr(), v1.data_ptr(), 1, 50, stream.cuda_stream), 5)
    out["truth_partner"] = {"workload": "one 64x64 state, 50 ticks in one launch (rollout's solver)",
                            "us_per_tick": t / 50 * 1e6, "cell_ticks_per_s": n * 50 / t}
    s.close()
    # large grid, tiled
    pl = sw.Params(nx=4096, ny=4096, nu=1e-3)
    sl = sw.ShallowWater(pl)
    hl, ul, vl = state(pl, 1, 2)
    steps = 20
    t = timed(lambda: sl.steps_dev(hl.data_ptr(), ul.data_ptr(), vl.data_ptr(), 1, steps, stream.cuda_stream), 2)
    nl = pl.cells()
    per_step = t / steps
    # the same bytes moved by a plain device copy (3 fields in, 3 out): the
    # practical ceiling of a step that reads and writes each cell once
    cin = torch.empty(3 * nl, dtype=torch.float32, device=dev).fill_(1.0)
    cout = torch.empty_like(cin)
    tc = timed(lambda: cout.copy_(cin), 5)
    del cin, cout
    out["large_grid"] = {"workload": f"one 4096x4096 state, {steps} steps, one HBM launch per step (row march)",
                         "kernel": "k_sw_band", "us_per_step": per_step * 1e6, "cell_steps_per_s": nl / per_step,
                         "algorithmic_bytes_per_step": 24 * nl, "achieved_GBps": 24 * nl / per_step / 1e9,
                         "frac_hbm": 24 * nl / per_step / PEAK["hbm"],
                         "copy_same_bytes_GBps": 24 * nl / tc / 1e9, "vs_copy": tc / per_step}
    sl.close()
    del hl, ul, vl
    # CPU: the oracle's step, one core, 64x64
    if seconds_cpu > 0:
        op = O.sw_params(nx=64, ny=64, nu=1e-3)
        hh = np.full(n, 1.0, np.float32)
        O.sw_add_drop(op, hh, 32.0, 32.0, 0.05, 4.0)
        z = np.zeros(n, np.float32)
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < seconds_cpu:
            hh, uu, vv = O.sw_steps(op, hh, z, z, 200)
            done += 200
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n * done / el, "unit": "cell-steps/s", "cores": 1, "kind": "port",
                               "sample": f"{done} steps of one 64x64 state, oracle restatement of "
                                         f"ShallowWater::step, {el:.1f} s"}
        out["dataset"]["vs_cpu_1core"] = out["dataset"]["cell_ticks_per_s"] / out["cpu_baseline"]["value"]
    return out


def run_scores(batch, grids, seconds_cpu):
    """eval-metrics' scores (metrics.rs via eval.rs:158-172) of `batch`
    64x64 trajectories in one j2k_score_dev call, and the excess-energy band
    (training project.py:66-140) over `grids` 64x64 grids in one
    j2k_project_excess_energy_dev call, with HIP events on a stream of their
    own; the CPU leg times the oracle's restatements on one core."""
    import ctypes

    import torch

    from jacobi2000_amd import checkpoint as ck
    from jacobi2000_amd._lib import Score, call
    import jacobi2000_amd as j
    from oracle import oracle as O

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    st, cj = ck.random_patch2d(seed=0)
    cfg = json.loads(cj)
    cfg["norm"] = {"mean": [0.0008, 0.0004, 0.0005], "std": [0.0063, 0.0138, 0.0135], "depth": 1.0}
    m = j.Model.from_bytes(st, json.dumps(cfg))
    n = 64 * 64
    rng = np.random.default_rng(0)
    out = {}
    with torch.cuda.stream(stream):
        x = torch.from_numpy(rng.standard_normal((batch, 3, n)).astype(np.float32)).to(dev)
        h = torch.from_numpy((1.0 + 0.01 * rng.standard_normal((batch, n))).astype(np.float32)).to(dev)
        u = torch.from_numpy((0.01 * rng.standard_normal((batch, n))).astype(np.float32)).to(dev)
        v = torch.from_numpy((0.01 * rng.standard_normal((batch, n))).astype(np.float32)).to(dev)
        centers = np.full((batch, 2), 32.0, np.float32)
        res = (Score * batch)()
        P = ctypes.c_void_p

        def score():
            call("j2k_score_dev", m.handle, batch, P(x.data_ptr()), P(h.data_ptr()), P(u.data_ptr()),
                 P(v.data_ptr()), centers.ctypes.data_as(P), 8, res, P(stream.cuda_stream))

        score()
        t0 = time.perf_counter()
        reps = 20
        for _ in range(reps):
            score()  # each call synchronises (the lag is finished on the host)
        t = (time.perf_counter() - t0) / reps
        out["scores"] = {"workload": f"{batch} trajectories of 64x64, one call (f64 sums on the device, lag on the "
                                     "host)", "ms_per_call": t * 1e3, "trajectories_per_s": batch / t}
        xg = torch.from_numpy(rng.standard_normal((grids, 3, n)).astype(np.float32)).to(dev)
        e = torch.zeros(grids, dtype=torch.float64, device=dev)
        call("j2k_excess_energy_dev", m.handle, grids, P(xg.data_ptr()), 0.0903, P(e.data_ptr()), P(stream.cuda_stream))
        prev = e * 0.5
        work = xg.clone()

        def project():
            work.copy_(xg)
            call("j2k_project_excess_energy_dev", m.handle, grids, P(work.data_ptr()), P(prev.data_ptr()), 0.0903,
                 0.9999, 1.0004, P(stream.cuda_stream))

        project()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(10):
            work.copy_(xg)
        b.record(stream)
        torch.cuda.synchronize()
        tcopy = a.elapsed_time(b) / 10 / 1e3
        a.record(stream)
        for _ in range(10):
            project()
        b.record(stream)
        torch.cuda.synchronize()
        tp = a.elapsed_time(b) / 10 / 1e3 - tcopy
        byt = grids * n * 12 * 2  # every grid is outside its band: read twice-ish, written once
        out["excess_band"] = {"workload": f"{grids} grids of 64x64 projected (all outside the band), one call",
                              "kernel": "k_excess<true>", "us_per_call": tp * 1e6, "grids_per_s": grids / tp,
                              "bytes_per_call": byt, "achieved_GBps": byt / tp / 1e9}
    if seconds_cpu > 0:
        mean, std, depth = cfg["norm"]["mean"], cfg["norm"]["std"], cfg["norm"]["depth"]
        xc, hc, uc, vc = (t_.cpu().numpy() for t_ in (x, h, u, v))
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < seconds_cpu:
            b_ = done % batch
            O.eval_scores(xc[b_], hc[b_], uc[b_], vc[b_], 64, 64, mean, std, depth, 32.0, 32.0, 8)
            done += 1
        el = time.perf_counter() - t0
        out["scores"]["cpu_baseline"] = {"value": done / el, "unit": "trajectories/s", "cores": 1, "kind": "port",
                                         "sample": f"{done} trajectories, oracle restatement of metrics.rs, "
                                                   f"{el:.1f} s"}
    return out


def stream_requests(n, seed=0):
    """Mixed request stream: sizes uniform over {32,48,64,96}^2 x {8,16}
    slices, mask kinds cycling forward / rewind / reconstruction (2% of cells),
    seeded random drops."""
    from jacobi2000_amd import checkpoint as ck

    rng = np.random.default_rng(seed)
    kinds = ("forward", "rewind", "reconstruction")
    reqs = []
    for i in range(n):
        side = int(rng.choice([32, 48, 64, 96]))
        T = int(rng.choice([8, 16]))
        g = ck.drops_grid(T, side, side, 3, seed=1000 + i)
        k = ck.mask_of(kinds[i % 3], T, side, side, seed=2000 + i)
        reqs.append((g, k))
    return reqs


def run_stream(st, cj, meta, reqs, rate, served, precision="bf16x2"):
    """Poisson arrivals at `rate` requests/s. served: the server in the given
    low-precision form with continuous batching; otherwise the f32 plain
    form, one request at a time in arrival order. Returns throughput and
    per-request latency."""
    import ctypes
    import jacobi2000_amd as j
    from jacobi2000_amd import _lib

    rng = np.random.default_rng(7)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
    cells = sum(g[0].size for g, _ in reqs)
    lat, iters, conv = [], [], 0
    if served:
        model = j.Model.from_bytes(st, cj, precision=precision)
        h = ctypes.c_void_p()
        _lib.call("j2k_server_create", model.handle, 64 << 30, 256, ctypes.byref(h))
        ids = []
        t0 = time.perf_counter()
        for (g, k), ta in zip(reqs, arrivals):
            wait = t0 + ta - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            rid = ctypes.c_uint64()
            _lib.call("j2k_server_submit", h, g.ctypes.data_as(ctypes.c_void_p), k.ctypes.data_as(ctypes.c_void_p),
                      meta.ctypes.data_as(ctypes.c_void_p), g.shape[1], g.shape[2], g.shape[3], ctypes.byref(rid))
            ids.append(rid.value)
        _lib.call("j2k_server_drain", h)
        first = last = None
        for rid in ids:
            r = _lib.Result()
            _lib.call("j2k_server_poll", h, rid, ctypes.byref(r), None)
            lat.append(r.done_s - r.submit_s)
            iters.append(r.iters)
            conv += r.converged
            first = r.submit_s if first is None else min(first, r.submit_s)
            last = r.done_s if last is None else max(last, r.done_s)
        launches, replays, peak = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
        _lib.call("j2k_server_stats", h, ctypes.byref(launches), ctypes.byref(replays), ctypes.byref(peak), None)
        _lib.call("j2k_server_destroy", h)
        model.close()
        span = last - first
        extra = {"iterations_launched": launches.value, "graph_replays": replays.value, "peak_active": peak.value}
    else:
        model = j.Model.from_bytes(st, cj, precision="f32")
        t0 = time.perf_counter()
        for (g, k), ta in zip(reqs, arrivals):
            wait = t0 + ta - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            _, it, r, v = model.iterate(j.fill_blank(g, k), k, meta)
            lat.append(time.perf_counter() - (t0 + ta))
            iters.append(it)
            conv += int(r < 1e-4)
        span = time.perf_counter() - t0
        model.close()
        extra = {}
    lat_ms = np.array(lat) * 1e3
    return dict(value=cells / span, seconds=span, requests=len(reqs), cells=cells, converged=conv,
                capped=len(reqs) - conv, mean_iterations=float(np.mean(iters)),
                p50_ms=float(np.percentile(lat_ms, 50)), p99_ms=float(np.percentile(lat_ms, 99)), **extra)


def run_wide(meta, steps, err_side=32):
    """The wide shape (d=256, 6 layers, 8 heads x 32, MLP 1024, seed 1, 5x5x3
    stencil) on 128x128 x 16 grids: f32, bf16x2 and FP8 throughput (one
    request per step, fixed point to tol), and each form's error against the
    f64 oracle on a err_side x err_side x 16 grid (the oracle at full size
    would take minutes), same iteration count as the GPU run there."""
    import torch
    import jacobi2000_amd as j
    from jacobi2000_amd import checkpoint as ck
    from oracle import oracle as O

    st, cj = ck.random_cell_st(dim=256, layers=6, heads=8, ff=1024, seed=1, radius=2, time_radius=1)
    T_, side = 16, 128
    dev = torch.device("cuda:0")
    out = {"workload": "wide shape d256 L6 H8x32 ff1024 seed 1, 5x5x3 stencil (75 keys), 128x128 x 16, "
                       "5 drops, slice 0 known, tol 1e-4, cap 32",
           "params": 0, "forms": {}}
    g = ck.drops_grid(T_, side, side, 5, seed=11)
    k = ck.mask_of("forward", T_, side, side)
    x0 = j.fill_blank(g, k)
    ge = ck.drops_grid(T_, err_side, err_side, 2, seed=12)
    ke = ck.mask_of("forward", T_, err_side, err_side)
    xe = j.fill_blank(ge, ke)
    cpu = O.FormC.load(st, cj)
    out["params"] = int(sum(a.size for a in cpu.w.values()))
    refs = {}
    for prec in ("f32", "bf16x2", "fp8"):
        model = j.Model.from_bytes(st, cj, precision=prec)
        d_x0 = torch.from_numpy(x0).to(dev)
        d_grid = torch.empty_like(d_x0)
        d_k = torch.from_numpy(k).to(dev)
        d_m = torch.from_numpy(meta).to(dev)
        stream = torch.cuda.current_stream().cuda_stream

        def one():
            d_grid.copy_(d_x0)
            return model.iterate_dev(d_grid.data_ptr(), d_k.data_ptr(), d_m.data_ptr(), T_, side, side, 32, 1e-4,
                                     stream)

        one()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its = []
        for _ in range(steps):
            its.append(one()[0])
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / steps
        xg, ite, _, ve = model.iterate(xe, ke, meta, max_iter=32, tol=1e-4)
        if ite not in refs:
            refs[ite] = cpu.iterate(xe, ke, meta, ite, 0.0, f64=True)[0]
        ref = refs[ite]
        d = (xg.astype(np.float64) - ref).reshape(3, -1)
        r = ref.reshape(3, -1)
        out["forms"][prec] = {
            "ms_per_request": el * 1e3, "cells_per_s": T_ * side * side / el, "iterations": its,
            "error_grid": f"{err_side}x{err_side}x{T_}", "error_iterations": ite,
            "rel_l2_vs_f64": float(np.linalg.norm(d) / np.linalg.norm(r)),
            "max_abs_per_field": [float(np.abs(d[c]).max()) for c in range(3)], "energy_violation": float(ve)}
        model.close()
    return out


def run_rollouts(st, cj, meta, n, served, precision="bf16x2", ny=96, nx=96, window=8, ticks=50):
    """Long rollouts: n forward requests on ny x nx grids, each chaining
    `window`-slice windows to `ticks` slices. served: all submitted at once to
    the server (chained on the device side, windows of different requests
    packed together); otherwise the f32 plain form runs every window of every
    rollout one at a time. Per-window latency and energy violation."""
    import ctypes
    import jacobi2000_amd as j
    from jacobi2000_amd import _lib
    from jacobi2000_amd import checkpoint as ck

    states = [ck.drops_grid(1, ny, nx, 5, seed=3000 + i)[:, 0].copy() for i in range(n)]
    lat, viol, iters = [], [], []
    if served:
        model = j.Model.from_bytes(st, cj, precision=precision)
        h = ctypes.c_void_p()
        _lib.call("j2k_server_create", model.handle, 128 << 30, 4096, ctypes.byref(h))
        t0 = time.perf_counter()
        ids = []
        for s0 in states:
            rid = ctypes.c_uint64()
            _lib.call("j2k_server_submit_rollout", h, s0.ctypes.data_as(ctypes.c_void_p),
                      meta.ctypes.data_as(ctypes.c_void_p), ny, nx, window, ticks, ctypes.byref(rid))
            ids.append(rid.value)
        _lib.call("j2k_server_drain", h)
        span = time.perf_counter() - t0
        for rid in ids:
            cnt = ctypes.c_int()
            la, vi, it = (ctypes.c_double * 64)(), (ctypes.c_double * 64)(), (ctypes.c_int * 64)()
            _lib.call("j2k_server_windows", h, rid, 64, ctypes.byref(cnt), la, vi, it)
            lat += list(la)[:cnt.value]
            viol += list(vi)[:cnt.value]
            iters += list(it)[:cnt.value]
        peak = ctypes.c_int()
        _lib.call("j2k_server_stats", h, None, None, ctypes.byref(peak), None)
        _lib.call("j2k_server_destroy", h)
        model.close()
        extra = {"peak_active": peak.value}
    else:
        model = j.Model.from_bytes(st, cj, precision="f32")
        t0 = time.perf_counter()
        for s0 in states:
            anchor, tt = s0, 0
            while tt < ticks - 1:
                T = min(window, ticks - tt)
                g = np.ascontiguousarray(np.repeat(anchor[:, None], T, axis=1))
                k = np.zeros((T, ny, nx), np.uint8)
                k[0] = 1
                ts = time.perf_counter()
                x, it, _, v = model.iterate(g, k, meta)
                lat.append(time.perf_counter() - ts)
                viol.append(v)
                iters.append(it)
                anchor, tt = x[:, T - 1].copy(), tt + window - 1
        span = time.perf_counter() - t0
        model.close()
        extra = {}
    cells = n * ticks * ny * nx
    lat_ms = np.array(lat) * 1e3
    return dict(value=cells / span, seconds=span, rollouts=n, windows=len(lat), cells=cells,
                mean_window_iterations=float(np.mean(iters)), max_energy_violation=float(max(viol)),
                window_p50_ms=float(np.percentile(lat_ms, 50)), window_p99_ms=float(np.percentile(lat_ms, 99)),
                **extra)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--stream", type=int, default=512,
                    help="also run the mixed request stream with this many requests (0: skip)")
    ap.add_argument("--rate", type=float, default=2000.0, help="Poisson arrival rate of the stream, requests/s")
    ap.add_argument("--wide", type=int, default=2,
                    help="timed requests per form of the wide shape on 128x128x16 grids (0: skip)")
    ap.add_argument("--rollouts", type=int, default=256,
                    help="also run this many chained 96x96 rollouts (8-slice windows to 50 ticks; 0: skip)")
    ap.add_argument("--headline", default="bf16x2", choices=["f32", "bf16", "bf16x2"],
                    help="served form reported as `value`")
    ap.add_argument("--no-oracle-error", action="store_true", help="skip the f64-oracle error check")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="after the runs, copy a buffer of this many MiB once (a known-size kernel that scales "
                         "the profiler's HBM byte counters)")
    ap.add_argument("--forms", default="f32,bf16,bf16x2",
                    help="forms to run (comma separated; f32 and the headline are always included)")
    ap.add_argument("--no-plain", action="store_true", help="skip the f32 plain form (profiling runs)")
    ap.add_argument("--solver", type=int, default=2048, metavar="STATES",
                    help="also time the device shallow-water solver with this many 64x64 dataset states (0: skip)")
    ap.add_argument("--scores", type=int, default=64, metavar="TRAJ",
                    help="also time the device scores on this many 64x64 trajectories and the excess band on 16x "
                         "as many grids (0: skip)")
    ap.add_argument("--iterating", type=int, default=10, metavar="STEPS",
                    help="timed requests per form of the iterating workload (input-sensitive checkpoint; 0: skip)")
    ap.add_argument("--form-p", type=int, default=20, metavar="STEPS",
                    help="timed steps per form of the reference forward at model20 shape (0: skip)")
    ap.add_argument("--no-cpu-legs", dest="cpu_legs", action="store_false",
                    help="skip the CPU oracle legs on 32x32x8 and 64x64x16 at all threads and one")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc runs that measure the dominant kernel's HBM bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one, and this path has no multi-GPU form "
              "(requests shard across GPUs without communication); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    st, cj, x0, known, meta = workload()
    runs, outs = {}, {}
    forms = [p for p in ("f32", "bf16", "bf16x2")
             if p in args.forms.split(",") or p == args.headline or (p == "f32" and not args.no_plain)]
    for prec in forms:
        note(f"form {prec}: {args.warmup} warmup + {args.steps} timed requests")
        runs[prec], outs[prec] = run_precision(prec, st, cj, x0, known, meta, args.steps, args.warmup)
    if args.copy_probe > 0:
        import torch

        a = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    # error of every form against the f64 oracle on this request (outside the timed region)
    errors = {}
    if not args.no_oracle_error:
        from oracle import oracle as O

        O.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        cpu = O.FormC.load(st, cj)
        for prec, out in outs.items():
            ref, _, _ = cpu.iterate(x0, known, meta, runs[prec]["iterations"][-1], 0.0, f64=True)
            d = (out.astype(np.float64) - ref).reshape(3, -1)
            r = ref.reshape(3, -1)
            errors[prec] = {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(r)),
                            "rel_l2_per_field": [float(np.linalg.norm(d[c]) / max(np.linalg.norm(r[c]), 1e-30))
                                                 for c in range(3)],
                            "max_abs_per_field": [float(np.abs(d[c]).max()) for c in range(3)]}
    head, plain = runs[args.headline], runs.get("f32", runs[args.headline])
    traffic = None
    if not args.no_traffic:
        k = head["kernels"]
        dom = max(k, key=lambda n: k[n]["ms_total"])
        sym = kernel_symbol(dom, args.headline, rows_pick_of(head))
        note(f"traffic: two rocprofv3 --pmc runs for {sym}")
        t_bytes, t_detail = measure_traffic(args, sym, f"rows_fused={rows_pick_of(head)}")
        traffic = t_bytes
        traffic_detail = t_detail
    summary = lambda r: {k: r[k] for k in ("value", "ms_per_step", "p50_ms", "p99_ms", "iterations", "resid",
                                          "energy_violation", "eager_timed_ms_per_step")}
    line = {
        "metric": "converged grid-cells/s + p50/p99 latency, BF16 served vs FP32 plain",
        "value": head["value"],
        "unit": "converged grid-cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.headline,
        "data": "synthetic: seeded random-init weights N(0,0.02), seeded Gaussian drops",
        "config": {"workload": "cell_st fixed point, one request 64x64 x 16 slices, 27-key stencil, "
                               "model20 shape (d128 L4 H4x32 ff512), slice 0 known, tol 1e-4, cap 32",
                   "cells_per_request": CELLS},
        "p50_ms": head["p50_ms"],
        "p99_ms": head["p99_ms"],
        "iterations": head["iterations"],
        "energy_violation": head["energy_violation"],
        "forms": {p: summary(r) for p, r in runs.items()},
        "vs_fp32_plain": {p: r["value"] / plain["value"] for p, r in runs.items()},
        "error_vs_f64_oracle": errors,
        "kernels": {p: r["kernels"] for p, r in runs.items()},
        "roofline": roofline(head, args.headline, traffic),
        "roofline_by_form": {p: roofline(r, p) for p, r in runs.items()},
        "autotune": head.get("autotune"),
    }
    if not args.no_traffic:
        line["traffic_detail"] = traffic_detail
    if args.stream > 0:
        note(f"stream: {args.stream} requests, served then plain")
        reqs = stream_requests(args.stream)
        s_served = run_stream(st, cj, meta, reqs, args.rate, served=True,
                              precision=args.headline if args.headline != "f32" else "bf16x2")
        s_plain = run_stream(st, cj, meta, reqs, args.rate, served=False)
        note(f"stream at saturation: {args.stream} requests all at t = 0, served then plain")
        z_served = run_stream(st, cj, meta, reqs, float("inf"), served=True,
                              precision=args.headline if args.headline != "f32" else "bf16x2")
        z_plain = run_stream(st, cj, meta, reqs, float("inf"), served=False)
        line["stream"] = {"workload": f"{args.stream} requests, Poisson arrivals at {args.rate:g}/s, sizes "
                                      "{32,48,64,96}^2 x {8,16}, masks forward/rewind/reconstruction",
                          "served": s_served, "plain_fp32": s_plain,
                          "served_vs_plain": s_served["value"] / s_plain["value"],
                          "saturated": {"workload": "the same requests, all submitted at t = 0 (closed loop: the "
                                                    "server's capacity)",
                                        "served": z_served, "plain_fp32": z_plain,
                                        "served_vs_plain": z_served["value"] / z_plain["value"]}}
    if args.wide > 0:
        note(f"wide shape: 128x128x16, f32 / bf16x2 / fp8, {args.wide} timed requests each, errors on 32x32x16")
        line["wide"] = run_wide(meta, args.wide)
    if args.rollouts > 0:
        note(f"rollouts: {args.rollouts} x 96x96, 8-slice windows to 50 ticks, served then plain")
        r_served = run_rollouts(st, cj, meta, args.rollouts, served=True,
                                precision=args.headline if args.headline != "f32" else "bf16x2")
        r_plain = run_rollouts(st, cj, meta, args.rollouts, served=False)
        line["rollouts"] = {"workload": f"{args.rollouts} concurrent forward rollouts, 96x96 grids, 8-slice windows "
                                        "chained to 50 ticks, energy violation checked per window",
                            "unit": "grid-cells/s over the whole trajectories",
                            "served": r_served, "plain_fp32": r_plain,
                            "served_vs_plain": r_served["value"] / r_plain["value"]}
    if args.iterating > 0:
        note(f"iterating workload: input-sensitive checkpoint, {args.iterating} timed requests per form")
        line["iterating"] = run_iterating(meta, args.iterating, args.warmup, args.headline, args.no_oracle_error)
    if args.form_p > 0:
        note(f"form P: reference forward at model20 shape, {args.form_p} timed steps per form")
        line["form_p"] = run_form_p(args.form_p, args.warmup, 0.0 if args.no_cpu_baseline else 2.0)
    if not args.no_cpu_baseline and args.cpu_legs:
        note("cpu legs: the oracle on 32x32x8 and 64x64x16, all host threads and one")
        line["cpu_baseline_legs"] = cpu_legs(st, cj, meta, x0, known)
    if args.solver > 0:
        note(f"solver: {args.solver} dataset states, truth partner, 4096x4096 grid")
        line["solver"] = run_solver(args.solver, 0.0 if args.no_cpu_baseline else 3.0)
    if args.scores > 0:
        note(f"scores: {args.scores} trajectories, excess band on {16 * args.scores} grids")
        line.update(run_scores(args.scores, 16 * args.scores, 0.0 if args.no_cpu_baseline else 2.0))
    if not args.no_cpu_baseline:
        note(f"cpu baseline: about {args.cpu_seconds:g} s of the CPU oracle")
        line["cpu_baseline"] = cpu_baseline(st, cj, x0, known, meta, max(plain["iterations"][0], 1), args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
