"""GPU parity of the cell_st form (per-cell space-time tokens, fixed point)
through the C ABI, against the CPU oracle: the model evaluation, every
iteration of the fixed point, the energy bound, bf16 error against the f64
oracle, and the continuous-batching server (each result the same alone and
in the mix)."""

import ctypes
import os
import time

import numpy as np
import pytest
import torch

import jacobi2000_amd as j
from jacobi2000_amd import _lib
from jacobi2000_amd import checkpoint as ck
from oracle import oracle as O

pytestmark = pytest.mark.gpu

META = ck.meta_of(1.0, 0.09, 9.81, 1.0, 1e-3)


@pytest.fixture(scope="module")
def small():
    return ck.random_cell_st(dim=32, layers=2, heads=2, ff=64, seed=3, std=0.1)


@pytest.fixture(scope="module")
def model20_cell():
    return ck.random_cell_st(seed=0)  # d=128, 4 layers, 4x32 heads, ff 512, 27-key stencil


def gpu_forward(model: j.Model, grid, known):
    _, T, ny, nx = grid.shape
    dev = torch.device("cuda:0")
    g = torch.from_numpy(np.ascontiguousarray(grid)).to(dev)
    k = torch.from_numpy(np.ascontiguousarray(known, dtype=np.uint8)).to(dev)
    m = torch.from_numpy(META).to(dev)
    y = torch.empty((T * ny * nx, 3), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    model.cell_forward_dev(g.data_ptr(), k.data_ptr(), m.data_ptr(), T, ny, nx, y.data_ptr(), stream)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("shape", [(4, 16, 16), (1, 3, 3), (3, 5, 7)])
def test_cell_forward_matches_oracle(small, shape):
    T, ny, nx = shape
    st, cj = small
    gpu = j.Model.from_bytes(st, cj)
    cpu = O.FormC.load(st, cj)
    rng = np.random.default_rng(sum(shape))
    g = rng.normal(0, 0.3, (3, T, ny, nx)).astype(np.float32)
    k = (rng.random((T, ny, nx)) < 0.3).astype(np.uint8)
    d = np.abs(gpu_forward(gpu, g, k) - cpu.forward(g, k, META)).max()
    print(f"cell forward {shape} max_abs={d:.3e}")
    assert d < 1e-5


def test_config1_forward_matches_oracle(model20_cell):
    """32x32 x 8 slices, model20 shape, 27-key stencil."""
    st, cj = model20_cell
    gpu = j.Model.from_bytes(st, cj)
    cpu = O.FormC.load(st, cj)
    g = ck.drops_grid(8, 32, 32, 1, 0, centred=True)
    k = ck.mask_of("forward", 8, 32, 32)
    x0 = O.fill_blank(g, k)
    y = gpu_forward(gpu, x0, k)
    ref = cpu.forward(x0, k, META)
    per_field = np.abs(y - ref).max(0)
    print("config1 forward max_abs per field", per_field)
    assert per_field.max() < 1e-4


@pytest.mark.parametrize("precision", ["f32", "bf16x2"])
def test_slices_above_the_register_path_match_oracle(small, precision):
    """Slices of more than 4096 cells take the projection's memory path (72 x
    72 = 5184 cells), 4096 or fewer its register path (64 x 64); every
    iteration against the oracle, the energy bound, the violation reported
    by the iteration itself."""
    st, cj = small
    gpu = j.Model.from_bytes(st, cj, precision=precision)
    cpu = O.FormC.load(st, cj)
    for T, n in ((4, 72), (4, 64)):
        g = ck.drops_grid(T, n, n, 2, 31)
        k = ck.mask_of("forward", T, n, n)
        x0 = O.fill_blank(g, k)
        xg, itg, rg, vg = gpu.iterate(x0, k, META, max_iter=2, tol=0.0)
        assert np.abs(xg - x0).max() > 1e-3  # the iterations move the blank slices
        if precision == "f32":
            xc, itc, _ = cpu.iterate(x0, k, META, 2, 0.0)
            d = np.abs(xg - xc).reshape(3, -1).max(1)
        else:
            xd, _, _ = cpu.iterate(x0, k, META, 2, 0.0, f64=True)
            d = np.linalg.norm(xg - xd) / np.linalg.norm(xd)
        print(f"{precision} {n}x{n}x{T}: {d} violation {vg:.2e}")
        assert itg == 2
        # bf16x2 keeps the weights in bf16: with this model's N(0, 0.1)
        # weights its increments carry about 1.5e-3 of the state against the
        # f64 oracle; a wrong projection or index shows as O(1)
        assert np.max(d) < (1e-4 if precision == "f32" else 5e-3)
        assert vg <= 1e-6


@pytest.mark.parametrize("kind", ["forward", "rewind", "reconstruction"])
def test_each_iteration_matches_oracle(small, kind):
    st, cj = small
    gpu = j.Model.from_bytes(st, cj)
    cpu = O.FormC.load(st, cj)
    T, ny, nx = 4, 12, 12
    g = ck.drops_grid(T, ny, nx, 2, 5)
    k = ck.mask_of(kind, T, ny, nx, seed=1, frac=0.1)
    x0 = O.fill_blank(g, k)
    for it in (1, 2, 3):
        xg, itg, rg, vg = gpu.iterate(x0, k, META, max_iter=it, tol=0.0)
        xc, itc, trc = cpu.iterate(x0, k, META, it, 0.0)
        per_field = np.abs(xg - xc).reshape(3, -1).max(1)
        print(f"{kind} iteration {it}: max_abs per field {per_field}, resid gpu={rg:.3e} cpu={trc[-1]:.3e}")
        assert itg == itc == it
        assert per_field.max() < 1e-4
        assert vg <= 1e-6


@pytest.mark.parametrize("kind", ["forward", "rewind", "reconstruction"])
def test_config1_fixed_point(model20_cell, kind):
    """Config 1: 32x32 x 8, drop h = 1 + 0.1 exp(-r^2/18), 8 iterations cap,
    projection on; per-field error per iteration against the f32 oracle and
    the energy bound."""
    st, cj = model20_cell
    gpu = j.Model.from_bytes(st, cj)
    cpu = O.FormC.load(st, cj)
    g = ck.drops_grid(8, 32, 32, 1, 0, centred=True)
    k = ck.mask_of(kind, 8, 32, 32, seed=1)
    x0 = O.fill_blank(g, k)
    xg, itg, rg, vg = gpu.iterate(x0, k, META, max_iter=8, tol=1e-4)
    xc, itc, trc = cpu.iterate(x0, k, META, 8, 1e-4)
    per_field = np.abs(xg - xc).reshape(3, -1).max(1)
    print(f"config1 {kind}: iters gpu={itg} cpu={itc} max_abs per field {per_field} violation {vg:.2e}")
    assert itg == itc
    assert per_field.max() < 1e-4
    assert vg <= 1e-6


@pytest.mark.parametrize("precision", ["bf16x2", "bf16"])
def test_served_forms_error_vs_f64_oracle(model20_cell, precision):
    """bf16x2 (bf16 weights, split-bf16 activations) converges like f32 and
    stays within 1e-3 relative L2 of the f64 oracle. Pure bf16 activations
    hold max|dx| near 1e-3, above tol, so the iteration is retired by the
    stall rule; its error is reported against a looser bound."""
    st, cj = model20_cell
    gpu = j.Model.from_bytes(st, cj, precision=precision)
    cpu = O.FormC.load(st, cj)
    T, ny, nx = 8, 32, 32
    g = ck.drops_grid(T, ny, nx, 5, 11)
    k = ck.mask_of("forward", T, ny, nx)
    x0 = O.fill_blank(g, k)
    xg, itg, rg, vg = gpu.iterate(x0, k, META, max_iter=32, tol=1e-4)
    xd, _, _ = cpu.iterate(x0, k, META, itg, 0.0, f64=True)
    rel = np.linalg.norm((xg - xd).reshape(3, -1), axis=1) / np.linalg.norm(xd.reshape(3, -1), axis=1).clip(1e-30)
    rel_all = np.linalg.norm(xg - xd) / np.linalg.norm(xd)
    print(f"{precision} iters={itg} resid={rg:.2e} rel L2 per field {rel} overall {rel_all:.3e} violation {vg:.2e}")
    assert vg <= 1e-6
    assert itg < 32
    if precision == "bf16x2":
        assert rg < 1e-4
        assert rel_all < 1e-3
    else:
        assert rel_all < 2e-2


@pytest.mark.parametrize("precision", ["f32", "bf16x2"])
def test_wide_model_5x5x3_stencil(precision):
    """The wide shape (d=256, 8 heads x 32, ff 1024, 75-key stencil, seed 1)
    on a small grid: every field within 1e-4 (f32) of the oracle per
    iteration; bf16x2 error against the f64 oracle reported and bounded."""
    st, cj = ck.random_cell_st(dim=256, layers=2, heads=8, ff=1024, seed=1, radius=2)
    gpu = j.Model.from_bytes(st, cj, precision=precision)
    cpu = O.FormC.load(st, cj)
    T, ny, nx = 4, 16, 16
    g = ck.drops_grid(T, ny, nx, 3, 21)
    k = ck.mask_of("forward", T, ny, nx)
    x0 = O.fill_blank(g, k)
    xg, itg, rg, vg = gpu.iterate(x0, k, META, max_iter=2, tol=0.0)
    if precision == "f32":
        xc, _, _ = cpu.iterate(x0, k, META, 2, 0.0)
        d = np.abs(xg - xc).reshape(3, -1).max(1)
        print(f"wide f32 max_abs per field {d}")
        assert d.max() < 1e-4
    else:
        xd, _, _ = cpu.iterate(x0, k, META, 2, 0.0, f64=True)
        rel = np.linalg.norm(xg - xd) / np.linalg.norm(xd)
        print(f"wide bf16x2 rel L2 {rel:.3e}")
        # The wide configuration states its error per precision (no target);
        # the bf16 weights of the 256-wide layers put it near 2e-3 here, and
        # an indexing error would show as O(1).
        assert rel < 5e-3
    assert vg <= 1e-6


def _model_env(st, cj, precision, **env):
    """A model created with the given environment (the tuning switches are
    read at creation)."""
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return j.Model.from_bytes(st, cj, precision=precision)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


SPECS = ["gemm_bm=64,attn_kv_mode=0,rows_fused=1", "gemm_bm=128,attn_kv_mode=1,rows_fused=0",
         "gemm_bm=64,attn_kv_mode=2,rows_fused=0", "gemm_bm=128,attn_kv_mode=2,rows_fused=2",
         "gemm_bm=64,attn_kv_mode=1,rows_fused=3", "gemm_bm=128,attn_kv_mode=0,rows_fused=4",
         "gemm_bm=64,attn_kv_mode=3,rows_fused=1", "gemm_bm=128,attn_kv_mode=4,rows_fused=0"]


@pytest.mark.parametrize("precision", ["f32", "bf16", "bf16x2"])
@pytest.mark.parametrize("spec", SPECS)
def test_launch_choices_do_not_change_results(model20_cell, precision, spec):
    """Every launch shape the autotuner may pick (GEMM block height, K/V
    staging mode of the attention, row-panel shape, the first block made
    from the embedding) gives bit-identical iterates, also on a grid whose
    token count is no multiple of the 64-token panel (3 x 20 x 22 = 1,320)."""
    st, cj = model20_cell
    base = _model_env(st, cj, precision, J2K_AUTOTUNE="0")
    forced = _model_env(st, cj, precision, J2K_TUNE_FORCE=spec)
    for T, ny, nx in ((8, 32, 32), (3, 20, 22)):
        g = ck.drops_grid(T, ny, nx, 3, 17)
        k = ck.mask_of("forward", T, ny, nx)
        x0 = O.fill_blank(g, k)
        xa, ia, ra, _ = base.iterate(x0, k, META, max_iter=2, tol=0.0)
        xb, ib, rb, _ = forced.iterate(x0, k, META, max_iter=2, tol=0.0)
        assert (ia, ra) == (ib, rb), (T, ny, nx)
        assert np.array_equal(xa, xb), (T, ny, nx)


@pytest.mark.parametrize("precision", ["f32", "bf16x2", "bf16"])
def test_wide_stencil_staging_modes_agree(precision):
    """The 75-key stencil in every staging mode gives the same bits: K and V
    together (every score in registers, a prefetch plus a batched
    remainder), K then V (scores grouped by four keys over the quad, two
    blocks per CU), one plane at a time, and two queries a lane; the time
    windows clipped at both ends (3 slices) and partial tiles (24 x 20)."""
    st, cj = ck.random_cell_st(dim=256, layers=1, heads=8, ff=1024, seed=1, radius=2)
    models = [_model_env(st, cj, precision, J2K_TUNE_FORCE=f"attn_kv_mode={m}") for m in range(5)]
    T, ny, nx = 3, 24, 20
    g = ck.drops_grid(T, ny, nx, 2, 23)
    k = ck.mask_of("rewind", T, ny, nx)
    x0 = O.fill_blank(g, k)
    xs = [m.iterate(x0, k, META, max_iter=1, tol=0.0)[0] for m in models]
    assert all(np.array_equal(xs[0], x) for x in xs[1:])


@pytest.mark.parametrize("heads", [2, 4])
@pytest.mark.parametrize("precision", ["f32", "bf16x2"])
def test_narrow_heads_wide_stencil(precision, heads):
    """The 75-key stencil with 16- and 8-dimension heads (d = 32, 2 or 4
    heads; scores grouped by four keys wherever two blocks share a CU):
    every staging mode gives the same bits, and f32 stays within 1e-4 of
    the oracle per field over two iterations (moving drops and r_max 1.05,
    so that h, u and v all depend on the model)."""
    st, cj = ck.random_cell_st(dim=32, layers=2, heads=heads, ff=64, seed=7, std=0.1, radius=2, r_max=1.05)
    models = [_model_env(st, cj, precision, J2K_TUNE_FORCE=f"attn_kv_mode={m}") for m in range(5)]
    T, ny, nx = 3, 24, 20
    g = ck.moving_drops_grid(T, ny, nx, 2, 29)
    k = ck.mask_of("forward", T, ny, nx)
    x0 = O.fill_blank(g, k)
    print(f"case narrow heads {precision} heads={heads}", flush=True)
    xs = [m.iterate(x0, k, META, max_iter=2, tol=0.0)[0] for m in models]
    assert all(np.array_equal(xs[0], x) for x in xs[1:])
    if precision == "f32":
        xc, _, _ = O.FormC.load(st, cj).iterate(x0, k, META, 2, 0.0)
        d = np.abs(xs[0] - xc).reshape(3, -1).max(1)
        print(f"narrow heads f32 max_abs per field {d}")
        assert d.max() < 1e-4


def test_autotune_table_and_results(model20_cell):
    """The tuner fills one entry per bucket with measured candidates, and a
    tuned model iterates exactly like an untuned one."""
    st, cj = model20_cell
    tuned = _model_env(st, cj, "bf16x2")
    rep = tuned.autotune(1 << 14)
    print("autotune", rep)
    assert rep["tuned_to_tokens"] == 1 << 14
    assert [b["tokens"] for b in rep["buckets"]] == [4096, 8192, 16384]
    for b in rep["buckets"]:
        assert b["gemm_qkv_bm"] in (64, 128) and b["gemm_proj_bm"] in (64, 128)
        assert b["attn_kv_mode"] in (0, 1, 2, 3, 4) and b["rows_fused"] in (0, 1, 2, 3, 4)
        assert min(b["gemm_qkv_bm_us"]) > 0 and min(b["attn_kv_mode_us"]) > 0
    plain = _model_env(st, cj, "bf16x2", J2K_AUTOTUNE="0")
    assert plain.autotune(0)["buckets"] == []
    T, ny, nx = 8, 48, 48
    g = ck.drops_grid(T, ny, nx, 3, 29)
    k = ck.mask_of("reconstruction", T, ny, nx, seed=4)
    x0 = O.fill_blank(g, k)
    xa = tuned.iterate(x0, k, META, max_iter=2, tol=0.0)[0]
    xb = plain.iterate(x0, k, META, max_iter=2, tol=0.0)[0]
    assert np.array_equal(xa, xb)


@pytest.mark.parametrize("precision", ["f32", "bf16x2"])
def test_iteration_graph_replays_match_eager_launches(model20_cell, precision):
    """From the second call with a grid shape on, j2k_iterate* replays each
    iteration from a HIP graph. Against a model that never captures
    (J2K_GRAPHS=0), every call must agree bit for bit (grid, iterations,
    residual, violation) across: repeats of one shape (capture, then
    replays), a larger shape (scratch grows, graphs retire), the first shape
    again, eager forwards on the same scratch between replays, a call with
    per-kernel timing on, and device calls on a stream of their own."""
    st, cj = model20_cell
    eager = _model_env(st, cj, precision, J2K_GRAPHS="0")
    graphs = _model_env(st, cj, precision)
    dev = torch.device("cuda:0")
    side = torch.cuda.Stream()
    reqs = []
    for i, (T, ny, nx, kind) in enumerate([(4, 16, 16, "forward"), (4, 16, 16, "rewind"), (4, 16, 16, "forward"),
                                          (8, 32, 32, "forward"), (8, 32, 32, "reconstruction"),
                                          (8, 32, 32, "forward"), (4, 16, 16, "rewind"), (4, 16, 16, "forward"),
                                          (4, 16, 16, "reconstruction")]):
        g = ck.drops_grid(T, ny, nx, 2, 40 + i)
        k = ck.mask_of(kind, T, ny, nx, seed=i)
        reqs.append((O.fill_blank(g, k), k))
    for i, (x0, k) in enumerate(reqs):
        _, T, ny, nx = x0.shape
        want = eager.iterate(x0, k, META, max_iter=6, tol=0.0)
        if i == 5:
            _lib.call("j2k_model_timing", graphs.handle, 1, None)
        if i in (2, 7):  # device form on a stream of its own
            g = torch.from_numpy(np.ascontiguousarray(x0)).to(dev)
            kd = torch.from_numpy(np.ascontiguousarray(k, dtype=np.uint8)).to(dev)
            md = torch.from_numpy(META).to(dev)
            torch.cuda.synchronize()
            it, r, v = graphs.iterate_dev(g.data_ptr(), kd.data_ptr(), md.data_ptr(), T, ny, nx, 6, 0.0,
                                          stream=side.cuda_stream)
            side.synchronize()
            got = (g.cpu().numpy(), it, r, v)
        else:
            got = graphs.iterate(x0, k, META, max_iter=6, tol=0.0)
        if i == 5:
            _lib.call("j2k_model_timing", graphs.handle, 0, None)
        assert got[1:] == want[1:], i
        assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32)), i
        if i in (1, 4, 6):  # eager work on the same scratch between replays
            y = gpu_forward(graphs, x0, k)
            assert np.array_equal(y.view(np.uint32), gpu_forward(eager, x0, k).view(np.uint32)), i


def test_device_iterations_replay_per_caller_buffers(model20_cell):
    """The device form's graphs hold the caller's grid, mask and meta
    pointers (the first iteration packs from them, every iteration unpacks
    into the grid). Alternating two sets of buffers of one shape, each call
    refilled with a new request, must agree bit for bit with a model that
    never captures: a change of buffers runs eagerly and captures again."""
    st, cj = model20_cell
    eager = _model_env(st, cj, "bf16x2", J2K_GRAPHS="0")
    graphs = _model_env(st, cj, "bf16x2")
    dev = torch.device("cuda:0")
    T, ny, nx = 4, 16, 16
    bufs = [(torch.empty((3, T, ny, nx), dtype=torch.float32, device=dev),
             torch.empty((T, ny, nx), dtype=torch.uint8, device=dev),
             torch.empty(META.shape, dtype=torch.float32, device=dev)) for _ in range(2)]
    for i, b in enumerate([0, 0, 0, 1, 1, 1, 0, 0]):
        g = ck.drops_grid(T, ny, nx, 2, 70 + i)
        k = ck.mask_of(("forward", "rewind", "reconstruction")[i % 3], T, ny, nx, seed=i)
        x0 = O.fill_blank(g, k)
        meta = (META * (1.0 + 0.1 * i)).astype(np.float32)
        want = eager.iterate(x0, k, meta, max_iter=3, tol=0.0)
        dg, dk, dm = bufs[b]
        dg.copy_(torch.from_numpy(np.ascontiguousarray(x0)))
        dk.copy_(torch.from_numpy(np.ascontiguousarray(k, dtype=np.uint8)))
        dm.copy_(torch.from_numpy(meta))
        torch.cuda.synchronize()
        it, r, v = graphs.iterate_dev(dg.data_ptr(), dk.data_ptr(), dm.data_ptr(), T, ny, nx, 3, 0.0)
        torch.cuda.synchronize()
        assert (it, r, v) == want[1:], i
        assert np.array_equal(dg.cpu().numpy().view(np.uint32), want[0].view(np.uint32)), i


@pytest.mark.parametrize("shape", ["model20", "wide"])
def test_fp8_form_error_vs_f64_oracle(model20_cell, shape):
    """The FP8 form (E4M3 block products, per-tensor weight and per-token
    activation scales, f32 accumulate)
    on a small grid: relative L2 against the f64 oracle after 2 iterations,
    reported per shape (no target is stated for FP8), and the energy bound."""
    st, cj = model20_cell if shape == "model20" else ck.random_cell_st(dim=256, layers=2, heads=8, ff=1024, seed=1,
                                                                          radius=2)
    gpu = j.Model.from_bytes(st, cj, precision="fp8")
    cpu = O.FormC.load(st, cj)
    T, ny, nx = 4, 24, 24
    g = ck.drops_grid(T, ny, nx, 3, 31)
    k = ck.mask_of("forward", T, ny, nx)
    x0 = O.fill_blank(g, k)
    xg, _, _, vg = gpu.iterate(x0, k, META, max_iter=2, tol=0.0)
    xd, _, _ = cpu.iterate(x0, k, META, 2, 0.0, f64=True)
    rel = np.linalg.norm(xg - xd) / np.linalg.norm(xd)
    inc = np.linalg.norm(xg - x0) / max(np.linalg.norm(xd - x0), 1e-30)
    print(f"fp8 {shape}: rel L2 {rel:.3e}, increment ratio {inc:.3f}")
    # The FP8 forward is off by about 3.5% (model20) and 2-4% (wide) against
    # the exact forward (E4M3 keeps 3 mantissa bits, with each token's own
    # scale). model20's fixed point moves the blank h cells by only ~3e-3
    # about the slice mean, below that per-token noise, so its relative error
    # here is about 0.33; the wide model's is about 0.43. No target is stated
    # for FP8: the bound only catches a blow-up; the arithmetic is pinned
    # against its restatement in the next test.
    assert rel < 0.6
    assert vg <= 1e-6


@pytest.mark.parametrize("shape", ["model20", "wide"])
def test_fp8_forward_matches_quantized_restatement(model20_cell, shape):
    """GPU FP8 forward of one block against the numpy restatement of the
    same quantized arithmetic (E4M3 codes, per-tensor weight scales,
    per-token input scales): it must sit far closer to it than to the
    unquantized forward. Differences come from f32 sums and the code flips
    they cause (a token's scale is its own f32 maximum, so a rounding
    difference there moves its whole row); with more blocks each flip feeds
    the next quantization, so one block is compared."""
    from tests import fp8_emulation as E

    st, cj = (ck.random_cell_st(dim=128, layers=1, heads=4, ff=512, seed=0) if shape == "model20" else
              ck.random_cell_st(dim=256, layers=1, heads=8, ff=1024, seed=1, radius=2))
    cpu = O.FormC.load(st, cj)
    gpu = j.Model.from_bytes(st, cj, precision="fp8")
    T, ny, nx = 3, 6, 6
    g = ck.drops_grid(T, ny, nx, 2, 35)
    k = ck.mask_of("forward", T, ny, nx)
    x0 = O.fill_blank(g, k)
    y = gpu_forward(gpu, x0, k).astype(np.float64)
    emu = E.cell_forward(cpu.w, cpu.cfg, x0, k, META, fp8=True)
    ref = E.cell_forward(cpu.w, cpu.cfg, x0, k, META, fp8=False)
    d_emu = np.linalg.norm(y - emu) / np.linalg.norm(ref)
    d_ref = np.linalg.norm(emu - ref) / np.linalg.norm(ref)
    print(f"fp8 {shape}: |gpu - quantized| {d_emu:.2e}, |quantized - exact| {d_ref:.2e} (relative L2)")
    assert d_emu < 0.2 * d_ref  # measured: 0.04x (d=128), 0.15x (wide)


def test_fp8_checkpoint_codes_match_load_time_quantization(model20_cell):
    """An F32 checkpoint quantized at load and the same checkpoint converted
    to F8_E4M3 codes beforehand give bit-identical FP8 iterates (the C++ and
    the Python encoders agree); an F32 model reading the converted file
    equals one reading its dequantized values as F32 (the F8 decode path)."""
    st, cj = model20_cell
    conv = ck.convert_checkpoint(st, "F8_E4M3")
    deq = ck.write_safetensors(ck.read_safetensors(conv))
    T, ny, nx = 4, 16, 16
    g = ck.drops_grid(T, ny, nx, 3, 33)
    k = ck.mask_of("forward", T, ny, nx)
    x0 = O.fill_blank(g, k)
    a = j.Model.from_bytes(st, cj, precision="fp8").iterate(x0, k, META, max_iter=2, tol=0.0)[0]
    b = j.Model.from_bytes(conv, cj, precision="fp8").iterate(x0, k, META, max_iter=2, tol=0.0)[0]
    assert np.array_equal(a, b)
    c = j.Model.from_bytes(conv, cj, precision="f32").iterate(x0, k, META, max_iter=2, tol=0.0)[0]
    d = j.Model.from_bytes(deq, cj, precision="f32").iterate(x0, k, META, max_iter=2, tol=0.0)[0]
    assert np.array_equal(c, d)
    bf = ck.convert_checkpoint(st, "BF16")
    e = j.Model.from_bytes(bf, cj, precision="f32").iterate(x0, k, META, max_iter=2, tol=0.0)[0]
    f = j.Model.from_bytes(ck.write_safetensors(ck.read_safetensors(bf)), cj, precision="f32").iterate(
        x0, k, META, max_iter=2, tol=0.0)[0]
    assert np.array_equal(e, f)


def test_grid_smaller_than_stencil_is_rejected(small):
    gpu = j.Model.from_bytes(*small)
    with pytest.raises(j.J2KError) as e:
        gpu.iterate(np.zeros((3, 2, 2, 2), np.float32), np.zeros((2, 2, 2), np.uint8), META)
    assert e.value.kind == "ShapeMismatch"


def _server(model: j.Model, budget: int, max_active: int):
    h = ctypes.c_void_p()
    _lib.call("j2k_server_create", model.handle, budget, max_active, ctypes.byref(h))
    return h


def _submit(h, grid, known):
    _, T, ny, nx = grid.shape
    rid = ctypes.c_uint64()
    _lib.call("j2k_server_submit", h, grid.ctypes.data_as(ctypes.c_void_p), known.ctypes.data_as(ctypes.c_void_p),
              META.ctypes.data_as(ctypes.c_void_p), T, ny, nx, ctypes.byref(rid))
    return rid.value


def _result(h, rid, shape):
    out = np.empty(shape, np.float32)
    res = _lib.Result()
    _lib.call("j2k_server_poll", h, rid, ctypes.byref(res), out.ctypes.data_as(ctypes.c_void_p))
    return res, out


def _requests():
    reqs = []
    sizes = [(4, 12, 12), (8, 16, 16), (4, 10, 14), (8, 12, 12), (4, 16, 16), (8, 10, 10)]
    kinds = ["forward", "rewind", "reconstruction"]
    for i, (T, ny, nx) in enumerate(sizes):
        g = ck.drops_grid(T, ny, nx, 2, 100 + i)
        k = ck.mask_of(kinds[i % 3], T, ny, nx, seed=i, frac=0.1)
        reqs.append((g, k))
    return reqs


@pytest.mark.parametrize("precision", ["f32", "bf16", "bf16x2", "fp8"])
def test_server_results_same_alone_and_in_the_mix(small, precision):
    model = j.Model.from_bytes(*small, precision=precision)
    reqs = _requests()
    alone = [model.iterate(j.fill_blank(g, k), k, META) for g, k in reqs]
    h = _server(model, 1 << 34, 16)
    try:
        ids = []
        for i, (g, k) in enumerate(reqs):
            ids.append(_submit(h, g, k))
            if i == 2:
                time.sleep(0.01)  # later requests arrive while the first ones run
        _lib.call("j2k_server_drain", h)
        for (g, k), rid, (xa, ita, ra, va) in zip(reqs, ids, alone):
            res, out = _result(h, rid, g.shape)
            d = np.abs(out - xa).max()
            print(f"request {rid}: iters mix={res.iters} alone={ita} max_abs={d:.2e} violation={res.violation:.1e}")
            assert res.iters == ita
            assert d <= 1e-6
            assert res.violation <= 1e-6
            assert res.done_s >= res.start_s >= res.submit_s
        launches, replays, peak = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
        _lib.call("j2k_server_stats", h, ctypes.byref(launches), ctypes.byref(replays), ctypes.byref(peak), None)
        print(f"server: {launches.value} iterations launched, {replays.value} graph replays, peak {peak.value} active")
        assert peak.value >= 2
    finally:
        _lib.call("j2k_server_destroy", h)


def _windows_alone(model, state, window, ticks):
    """A chained rollout run window by window through j2k_iterate."""
    _, ny, nx = state.shape
    traj = np.zeros((3, ticks, ny, nx), np.float32)
    traj[:, 0] = state
    anchor, t0, iters, viol = state, 0, [], []
    while t0 < ticks - 1:
        T = min(window, ticks - t0)
        g = np.ascontiguousarray(np.repeat(anchor[:, None], T, axis=1))
        k = np.zeros((T, ny, nx), np.uint8)
        k[0] = 1
        x, it, _, v = model.iterate(g, k, META)
        traj[:, t0 + 1:t0 + T] = x[:, 1:T]
        anchor, t0 = x[:, T - 1], t0 + window - 1
        iters.append(it)
        viol.append(v)
    return traj, iters, viol


@pytest.mark.parametrize("precision", ["f32", "bf16x2"])
def test_server_rollout_chains_windows(small, precision):
    """Chained rollouts (window 4, 9 ticks: windows of 4, 4 and 3 slices),
    three at once with plain requests in the mix, each equal to its windows
    run one by one, every window within the energy bound; the plain
    requests (retired in turns with and without a rollout window) equal to
    their runs alone."""
    model = j.Model.from_bytes(*small, precision=precision)
    ny, nx, W, ticks = 16, 12, 4, 9
    states = [ck.drops_grid(1, ny, nx, 2, 40 + i)[:, 0].copy() for i in range(3)]
    alone = [_windows_alone(model, st, W, ticks) for st in states]
    plain = _requests()[:3]
    plain_alone = [model.iterate(j.fill_blank(g, k), k, META) for g, k in plain]
    h = _server(model, 1 << 34, 16)
    try:
        ids, plain_ids = [], []
        for i, st in enumerate(states):
            rid = ctypes.c_uint64()
            _lib.call("j2k_server_submit_rollout", h, st.ctypes.data_as(ctypes.c_void_p),
                      META.ctypes.data_as(ctypes.c_void_p), ny, nx, W, ticks, ctypes.byref(rid))
            ids.append(rid.value)
            g, k = plain[i]
            plain_ids.append(_submit(h, g, k))
        _lib.call("j2k_server_drain", h)
        for (g, k), rid, (xa, ita, _, _) in zip(plain, plain_ids, plain_alone):
            res, out = _result(h, rid, g.shape)
            print(f"plain {rid}: iters mix={res.iters} alone={ita} max_abs={np.abs(out - xa).max():.2e}")
            assert res.iters == ita and np.abs(out - xa).max() <= 1e-6
        for rid, (traj, iters, viol) in zip(ids, alone):
            res, out = _result(h, rid, (3, ticks, ny, nx))
            n = ctypes.c_int()
            lat, vio, its = (ctypes.c_double * 8)(), (ctypes.c_double * 8)(), (ctypes.c_int * 8)()
            _lib.call("j2k_server_windows", h, rid, 8, ctypes.byref(n), lat, vio, its)
            d = np.abs(out - traj).max()
            print(f"rollout {rid}: windows {n.value} iters {list(its)[:n.value]} alone {iters} max_abs {d:.2e}")
            assert n.value == 3 and list(its)[:3] == iters
            assert d <= 1e-6
            assert res.iters == sum(iters) and res.violation <= 1e-6
            assert all(vio[w] <= 1e-6 and lat[w] > 0 for w in range(3))
        with pytest.raises(j.J2KError):
            _lib.call("j2k_server_windows", h, ids[0] + 1, 8, None, None, None, None)  # a plain request
    finally:
        _lib.call("j2k_server_destroy", h)


def test_server_budget_admits_one_at_a_time(small):
    model = j.Model.from_bytes(*small)
    reqs = _requests()[:3]
    # f32, d=32, ff=64: 1076 device bytes per token; the budget holds the
    # 2048-token request alone but no two of the three requests together
    h = _server(model, 2048 * 1076 + 100, 16)
    try:
        ids = [_submit(h, g, k) for g, k in reqs]
        _lib.call("j2k_server_drain", h)
        peak = ctypes.c_int()
        _lib.call("j2k_server_stats", h, None, None, ctypes.byref(peak), None)
        assert peak.value == 1
        for (g, k), rid in zip(reqs, ids):
            res, out = _result(h, rid, g.shape)
            xa, ita, _, _ = model.iterate(j.fill_blank(g, k), k, META)
            assert res.iters == ita and np.abs(out - xa).max() <= 1e-6
    finally:
        _lib.call("j2k_server_destroy", h)


def test_server_pending_and_unknown_ids(small):
    model = j.Model.from_bytes(*small)
    h = _server(model, 1 << 30, 4)
    try:
        res = _lib.Result()
        rc = _lib.lib().j2k_server_poll(h, 12345, ctypes.byref(res), None)
        assert rc == _lib.ERR_INVALID
        g, k = _requests()[0]
        rid = _submit(h, g, k)
        rc = _lib.lib().j2k_server_poll(h, rid, ctypes.byref(res), None)
        assert rc in (_lib.OK, _lib.PENDING)
        _lib.call("j2k_server_drain", h)
        assert _lib.lib().j2k_server_poll(h, rid, ctypes.byref(res), None) == _lib.OK
    finally:
        _lib.call("j2k_server_destroy", h)
