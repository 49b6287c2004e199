"""The allocation-free entry point (j2k_cell_forward_dev_noalloc) against
j2k_cell_forward_dev through the C ABI: bit-identical outputs on two calls
in a row, the second on a smaller grid, in the plain and the served forms;
the model's own path unchanged after it; a grid beyond the fixed region
refused before any device work."""

import ctypes

import numpy as np
import pytest
import torch

import jacobi2000_amd as j
from jacobi2000_amd import _lib
from jacobi2000_amd import checkpoint as ck
from tests import served_checks as S

pytestmark = pytest.mark.gpu

META = S.META


def _buffers(grid, known):
    dev = torch.device("cuda:0")
    g = torch.from_numpy(np.ascontiguousarray(grid)).to(dev)
    k = torch.from_numpy(np.ascontiguousarray(known, dtype=np.uint8)).to(dev)
    m = torch.from_numpy(META).to(dev)
    return g, k, m


def _forward(model, grid, known, noalloc):
    _, T, ny, nx = grid.shape
    g, k, m = _buffers(grid, known)
    y = torch.full((T * ny * nx, 3), float("nan"), dtype=torch.float32, device=g.device)
    torch.cuda.synchronize()
    name = "j2k_cell_forward_dev_noalloc" if noalloc else "j2k_cell_forward_dev"
    _lib.call(name, model.handle, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(k.data_ptr()),
              ctypes.c_void_p(m.data_ptr()), T, ny, nx, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(None))
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("precision", ["f32", "bf16x2", "bf16"])
def test_noalloc_forward_matches_forward(precision):
    st, cj = S.teeth()
    model = j.Model.from_bytes(st, cj, precision=precision)
    requests = [S.request("reconstruction", 16, 32, 32), S.request("forward", 8, 24, 24)]
    first = []
    for i, (x0, k) in enumerate(requests):
        print(f"case {precision} call {i + 1}: grid {x0.shape}", flush=True)
        want = _forward(model, x0, k, noalloc=False)
        got = _forward(model, x0, k, noalloc=True)
        assert np.isfinite(got).all()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), i
        first.append(want)
    # the model's own path after the allocation-free calls: its cached tables
    # and slots still give the same bits, for both grids
    for (x0, k), want in zip(requests, first):
        again = _forward(model, x0, k, noalloc=False)
        assert np.array_equal(again.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("pick", [0, 1, 2])
def test_noalloc_wide_token_local_launches(pick, monkeypatch):
    """d = 256 with each way of running the token-local work forced (tile
    GEMMs, the fused MLP, the whole-block launch): the allocation-free
    forward gives the model's own bits, on two calls in a row, the second
    on a smaller grid."""
    monkeypatch.setenv("J2K_TUNE_FORCE", f"rows_fused={pick}")
    st, cj = ck.random_cell_st(dim=256, layers=2, heads=8, ff=1024, seed=1, std=0.1, radius=2)
    model = j.Model.from_bytes(st, cj, precision="bf16x2")
    for i, (x0, k) in enumerate([S.request("reconstruction", 3, 24, 20), S.request("forward", 2, 16, 16)]):
        print(f"case wide rows_fused={pick} call {i + 1}: grid {x0.shape}", flush=True)
        want = _forward(model, x0, k, noalloc=False)
        got = _forward(model, x0, k, noalloc=True)
        assert np.isfinite(got).all()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), i


def test_noalloc_refuses_a_grid_beyond_the_region():
    st, cj = S.teeth()
    model = j.Model.from_bytes(st, cj, precision="f32")
    x0, k = S.request("forward", 16, 64, 64)  # 65,536 tokens: about 340 MB of f32 scratch
    with pytest.raises(j.J2KError) as e:
        _forward(model, x0, k, noalloc=True)
    print("refused:", e.value)
    assert e.value.code == _lib.ERR_OOM
    assert "fixed region" in str(e.value)
    # the refusal left the model usable, and the region too
    x1, k1 = S.request("forward", 4, 16, 16)
    a = _forward(model, x1, k1, noalloc=True)
    b = _forward(model, x1, k1, noalloc=False)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_noalloc_needs_a_cell_st_model():
    st, cj = ck.random_patch2d(nx=16, ny=16, dim=32, layers=1, heads=2, ff=64, head_dim=8)
    model = j.Model.from_bytes(st, cj)
    x0, k = S.request("forward", 2, 8, 8)
    with pytest.raises(j.J2KError):
        _forward(model, x0, k, noalloc=True)
