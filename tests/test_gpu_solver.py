"""The device shallow-water solver (libj2k j2k_sw_*) against the oracle's
restatement of crates/solvers/src/shallow_water.rs: bit-identical states
after any number of steps, on both kernels (LDS-resident and tiled), on
grids from 1x1 to 1024x768, band edges (60-column bands with two halo
lanes beyond 64 columns) and batches; dump-dataset
trajectories with drops; the reference's own solver tests run on the
device; host-side argument checks on the CPU."""

import ctypes

import numpy as np
import pytest

from jacobi2000_amd import _lib
from jacobi2000_amd import solvers as sw
from oracle import oracle as o

f32 = np.float32


def oparams(p: sw.Params) -> o.SwParams:
    return o.SwParams(p.nx, p.ny, p.dx, p.dt, p.g, p.depth, p.nu)


def seeded_state(p: sw.Params, seed: int, batch: int = 1):
    """Flat water plus a few drops and small random currents per state."""
    rng = np.random.default_rng(seed)
    n = p.cells()
    h = np.full((batch, n), p.depth, f32)
    u = (rng.standard_normal((batch, n)) * 1e-3).astype(f32)
    v = (rng.standard_normal((batch, n)) * 1e-3).astype(f32)
    for b in range(batch):
        for _ in range(3):
            f = sw.drop_field(p, rng.uniform(0, p.nx), rng.uniform(0, p.ny), rng.uniform(-0.1, 0.1),
                              rng.uniform(1.0, 5.0))
            h[b] += f.reshape(-1)
    return h, u, v


def oracle_steps(p, h, u, v, steps):
    out = [o.sw_steps(oparams(p), h[b], u[b], v[b], steps) for b in range(h.shape[0])]
    return tuple(np.stack([x[i] for x in out]) for i in range(3))


# ---- CPU: arguments and the host-side bump -------------------------------------

def test_drop_field_matches_the_oracle_bitwise():
    p = sw.Params(nx=37, ny=23)
    for cx, cy, amp, sig in [(20.0, 30.0, 0.05, 4.0), (3.3, 17.9, -0.2, 2.5), (0.0, 0.0, 0.08, 6.0)]:
        want = np.zeros(p.cells(), f32)
        o.sw_add_drop(oparams(p), want, cx, cy, amp, sig)
        assert (sw.drop_field(p, cx, cy, amp, sig).reshape(-1) == want).all()


def test_bad_grid_is_invalid_before_any_device_call():
    h = ctypes.c_void_p()
    for nx, ny in [(0, 4), (4, 0), (-1, 3), (65536, 32768)]:
        rc = _lib.lib().j2k_sw_create(ctypes.byref(_lib.SwParams(nx, ny, 1, 0.1, 9.81, 1, 0)), 0, ctypes.byref(h))
        assert rc == _lib.ERR_INVALID
        assert b"grid" in _lib.lib().j2k_last_error()


def test_params_mirror_the_reference_defaults():
    p = sw.Params()
    assert (p.nx, p.ny, p.dx, p.g, p.depth, p.nu) == (64, 64, 1.0, 9.81, 1.0, 0.0)
    assert p.dt == o.sw_params().dt
    # `Params { nu: 1e-3, ..Params::default() }` keeps the default dt
    assert sw.Params(nu=1e-3, g=5.0).dt == p.dt
    assert abs(p.cells_per_tick() - 0.4 / np.sqrt(2.0)) < 1e-6
    assert abs(p.horizon_ticks(20.0, 30.0) - 20.0 / p.cells_per_tick()) < 1e-4


# ---- GPU ---------------------------------------------------------------------------

def run_device(p, h, u, v, steps, kernel="auto"):
    s = sw.ShallowWater(p)
    s.set_kernel(kernel)
    st = sw.State(h.copy(), u.copy(), v.copy())
    s.steps(st, steps)
    s.close()
    return st.h, st.u, st.v


GRIDS = [(64, 64), (1, 1), (1, 9), (9, 1), (2, 2), (16, 16), (63, 5), (65, 17), (121, 3), (125, 7), (100, 68),
         (64, 106), (90, 90), (33, 130)]


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny", GRIDS)
@pytest.mark.parametrize("kernel", ["resident", "tiled"])
def test_steps_bit_identical_to_the_oracle(nx, ny, kernel):
    p = sw.Params(nx=nx, ny=ny, nu=1e-3)
    if kernel == "resident" and nx * ny > sw.ShallowWater.RESIDENT_MAX_CELLS:
        s = sw.ShallowWater(p)
        with pytest.raises(_lib.J2KError) as e:
            s.set_kernel("resident")
            s.steps(sw.State(*seeded_state(p, 0)), 1)
        assert e.value.code == _lib.ERR_LIMIT
        return
    h, u, v = seeded_state(p, nx * 1000 + ny, batch=2)
    for steps in (1, 2, 7):
        print(f"solver {kernel} {nx}x{ny} steps {steps}", flush=True)
        got = run_device(p, h, u, v, steps, kernel)
        want = oracle_steps(p, h, u, v, steps)
        for g, w in zip(got, want):
            assert (g.view(np.uint32) == w.view(np.uint32)).all(), (kernel, nx, ny, steps)


@pytest.mark.gpu
def test_other_physics_and_long_runs_bit_identical():
    p = sw.Params(nx=48, ny=40, dx=0.5, dt=0.03, g=3.7, depth=2.5, nu=0.02)
    h, u, v = seeded_state(p, 7, batch=3)
    h += f32(1.5)  # depth 2.5 around the drops
    for kernel in ("resident", "tiled"):
        got = run_device(p, h, u, v, 300, kernel)
        want = oracle_steps(p, h, u, v, 300)
        for g, w in zip(got, want):
            assert (g.view(np.uint32) == w.view(np.uint32)).all(), kernel


@pytest.mark.gpu
def test_large_grid_tiled_bit_identical():
    p = sw.Params(nx=1024, ny=768, nu=1e-3)
    h, u, v = seeded_state(p, 3)
    print("solver tiled 1024x768 x 4 steps", flush=True)
    got = run_device(p, h, u, v, 4)
    want = oracle_steps(p, h, u, v, 4)
    for g, w in zip(got, want):
        assert (g.view(np.uint32) == w.view(np.uint32)).all()


@pytest.mark.gpu
def test_zero_steps_and_empty_batch_are_no_ops():
    p = sw.Params(nx=20, ny=20)
    h, u, v = seeded_state(p, 1)
    got = run_device(p, h, u, v, 0)
    assert all((g == w).all() for g, w in zip(got, (h, u, v)))
    s = sw.ShallowWater(p)
    s.steps(sw.State(np.zeros(0, f32), np.zeros(0, f32), np.zeros(0, f32)), 5)
    with pytest.raises(_lib.J2KError) as e:
        s.steps(sw.State(h.copy(), u.copy(), v.copy()), -1)
    assert e.value.code == _lib.ERR_INVALID


@pytest.mark.gpu
def test_device_pointer_path_on_a_torch_stream():
    import torch

    p = sw.Params(nx=64, ny=64, nu=1e-3)
    h, u, v = seeded_state(p, 11, batch=4)
    dev = torch.device("cuda:0")
    th, tu, tv = (torch.from_numpy(a.copy()).to(dev) for a in (h, u, v))
    s = sw.ShallowWater(p)
    for kernel in ("resident", "tiled"):
        s.set_kernel(kernel)
        a, b, c = th.clone(), tu.clone(), tv.clone()
        s.steps_dev(a.data_ptr(), b.data_ptr(), c.data_ptr(), 4, 25, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        want = oracle_steps(p, h, u, v, 25)
        for g, w in zip((a, b, c), want):
            assert (g.cpu().numpy().view(np.uint32) == w.view(np.uint32)).all(), kernel


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["resident", "tiled"])
def test_trajectories_match_the_oracle_with_drops(kernel):
    p = sw.Params(nx=40, ny=28, nu=1e-3)
    B = 5
    h, u, v = seeded_state(p, 5, batch=B)
    ticks = [6, 0, 11, 3, 9]
    fields = np.stack([sw.drop_field(p, 10.0, 7.0, 0.04, 3.0), sw.drop_field(p, 30.0, 20.0, -0.03, 2.0),
                       sw.drop_field(p, 20.0, 14.0, 0.05, 5.0)])
    # (state, tick, field): two drops on one tick of state 2 (order matters),
    # one past state 3's last tick (never applied), one at a state's last tick
    drops = [(2, 4, 1), (0, 2, 0), (2, 4, 0), (2, 7, 2), (3, 5, 2), (4, 9, 1)]
    s = sw.ShallowWater(p)
    s.set_kernel(kernel)
    st = sw.State(h.copy(), u.copy(), v.copy())
    trajs = s.trajectories(st, ticks, drops, fields)
    for b in range(B):
        mine = [(tk, fi) for (sb, tk, fi) in drops if sb == b]
        want, hf, uf, vf = o.sw_trajectory(oparams(p), h[b], u[b], v[b], ticks[b], mine, fields)
        assert trajs[b].shape == (3, ticks[b] + 1, p.ny, p.nx)
        assert (trajs[b].view(np.uint32) == want.view(np.uint32)).all(), (kernel, b)
        for g, w in zip((st.h[b], st.u[b], st.v[b]), (hf, uf, vf)):
            assert (g.view(np.uint32) == w.view(np.uint32)).all(), (kernel, b)


# ---- the reference's solver tests on the device (shallow_water.rs:260-322) ------

@pytest.mark.gpu
def test_reference_solver_tests_on_device():
    p = sw.Params(nx=64, ny=64, nu=0.0)
    s = sw.ShallowWater(p)
    st = sw.State.flat(p)
    s.steps(st, 200)
    assert (st.h == f32(p.depth)).all() and not st.u.any() and not st.v.any()

    st = sw.State.flat(p)
    st.add_drop(p, 20.0, 30.0, 0.05, 4.0)
    m0 = st.mass()
    s.steps(st, 500)
    assert abs((st.mass() - m0) / m0) < 1e-7

    st = sw.State.flat(p)
    st.add_drop(p, 31.5, 31.5, 0.05, 5.0)
    s.steps(st, 120)
    h, u, v = (a.reshape(64, 64) for a in (st.h, st.u, st.v))
    assert np.abs(h - h[:, ::-1]).max() < 1e-6 and np.abs(h - h[::-1]).max() < 1e-6
    assert np.abs(u + u[:, ::-1]).max() < 1e-6 and np.abs(v + v[::-1]).max() < 1e-6

    p2 = sw.Params(nx=64, ny=64, nu=1e-3)
    s2 = sw.ShallowWater(p2)
    st = sw.State.flat(p2)
    st.add_drop(p2, 24.0, 40.0, 0.08, 4.0)
    s2.steps(st, 2000)
    mx = np.abs(st.h - f32(p2.depth)).max()
    assert np.isfinite(mx) and mx < 0.08
