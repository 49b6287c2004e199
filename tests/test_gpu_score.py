"""Device scoring (j2k_score*, j2k_correlation_dev) against the oracle's
restatement of crates/engine/src/metrics.rs as eval-metrics applies it
(crates/engine-cli/src/eval.rs:158-172), and the rollout CLI's loop with its
truth partner on the device (j2k_rollout_truth) against the device rollout
and the oracle's solver and correlation.

Tolerances: the reference sums correlation terms in f64 left to right and
the device in a fixed tree, so correlations agree to f64 rounding (2e-6
after the f32 cast); peaks are exact; mean and std_dev are f32 sums in the
reference and f64 here, so bias agrees to the reference's f32 summation
error (1e-4 relative); the lag correlates the same profiles (1e-4)."""

import ctypes

import numpy as np
import pytest

import jacobi2000_amd as j
from jacobi2000_amd import _lib
from jacobi2000_amd import solvers as sw
from oracle import oracle as O

f32 = np.float32
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tiny(fixture_dir):
    return j.Model.load(fixture_dir)


@pytest.fixture(scope="module")
def model20():
    from jacobi2000_amd import checkpoint as ck

    st, cj = ck.random_patch2d(seed=0)
    return j.Model.from_bytes(st, cj)


def norm_of(m):
    c = m.cfg
    return list(c.norm_mean), list(c.norm_std), c.norm_depth


def trajectories(m, B, seed):
    """Solver states and model-like states (the solver state encoded, then
    perturbed and shifted) for B drops on the model's grid."""
    c = m.cfg
    nx, ny = c.nx, c.ny
    rng = np.random.default_rng(seed)
    p = sw.Params(nx=nx, ny=ny, dx=c.dx, dt=c.dt, g=c.g, depth=c.depth, nu=c.nu)
    mean, std, depth = norm_of(m)
    H, U, V, X, C = [], [], [], [], []
    for b in range(B):
        cx, cy = rng.uniform(3, nx - 3), rng.uniform(3, ny - 3)
        st = sw.State.flat(p)
        st.add_drop(p, cx, cy, rng.uniform(0.03, 0.07), rng.uniform(1.5, 3.0))
        st.u += (rng.standard_normal(nx * ny) * 1e-3).astype(f32)
        st.v += (rng.standard_normal(nx * ny) * 1e-3).astype(f32)
        x = O.encode(st.h, st.u, st.v, mean, std, depth)
        x = (x * f32(rng.uniform(0.8, 1.2)) + (rng.standard_normal(x.size) * 0.02).astype(f32)).astype(f32)
        x[: nx * ny] = np.roll(x[: nx * ny].reshape(ny, nx), b % 3, axis=1).reshape(-1)
        H.append(st.h), U.append(st.u), V.append(st.v), X.append(x), C.append((cx, cy))
    return np.stack(X), np.stack(H), np.stack(U), np.stack(V), np.array(C, f32)


def check_scores(m, got, X, H, U, V, C):
    c = m.cfg
    mean, std, depth = norm_of(m)
    for b, g in enumerate(got):
        w = O.eval_scores(X[b], H[b], U[b], V[b], c.nx, c.ny, mean, std, depth, float(C[b, 0]), float(C[b, 1]), 8)
        assert abs(g["corr"] - w["corr"]) < 2e-6, (b, g, w)
        assert abs(g["corr_h"] - w["corr_h"]) < 2e-6, (b, g, w)
        assert g["amplitude_ratio"] == np.float32(w["amplitude_ratio"]), (b, g, w)
        assert abs(g["lag"] - w["lag"]) < 1e-4, (b, g, w)
        assert abs(g["bias"] - w["bias"]) < 1e-4 * max(1.0, abs(w["bias"])), (b, g, w)


def test_scores_match_the_oracle(tiny):
    X, H, U, V, C = trajectories(tiny, 6, 0)
    print("score: 6 trajectories on 16x16", flush=True)
    check_scores(tiny, tiny.score(X, H, U, V, C), X, H, U, V, C)


def test_scores_of_model20_grid(model20):
    X, H, U, V, C = trajectories(model20, 4, 1)
    print("score: 4 trajectories on 64x64", flush=True)
    check_scores(model20, model20.score(X, H, U, V, C), X, H, U, V, C)


def test_identical_states_score_one_and_zero(tiny):
    X, H, U, V, C = trajectories(tiny, 2, 2)
    mean, std, depth = norm_of(tiny)
    n = tiny.cfg.nx * tiny.cfg.ny
    X = np.stack([O.encode(H[b], U[b], V[b], mean, std, depth) for b in range(2)])
    for g in tiny.score(X, H, U, V, C):
        assert abs(g["corr"] - 1.0) < 1e-6 and abs(g["corr_h"] - 1.0) < 1e-6
        assert g["amplitude_ratio"] == 1.0 and g["bias"] == 0.0 and abs(g["lag"]) < 1e-6
    assert X.shape[1] == 3 * n


def test_correlation_dev_of_decoded_fields():
    import torch

    rng = np.random.default_rng(3)
    n = 4096
    a = rng.standard_normal(n).astype(f32)
    b = (a * f32(0.7) + rng.standard_normal(n).astype(f32) * f32(0.3)).astype(f32)
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    out = torch.zeros(1, dtype=torch.float32, device="cuda")
    for sa, ta, sb, tb in [(0.3, 0.1, 1.0, -1.0), (1.0, 0.0, 1.0, 0.0)]:
        _lib.call("j2k_correlation_dev", ctypes.c_void_p(da.data_ptr()), sa, ta, ctypes.c_void_p(db.data_ptr()), sb,
                  tb, n, ctypes.c_void_p(out.data_ptr()), None)
        torch.cuda.synchronize()
        want = O.correlation(a * f32(sa) + f32(ta), b * f32(sb) + f32(tb))
        assert abs(float(out.item()) - want) < 2e-6
    const = torch.ones(n, dtype=torch.float32, device="cuda")
    _lib.call("j2k_correlation_dev", ctypes.c_void_p(const.data_ptr()), 1.0, 0.0, ctypes.c_void_p(db.data_ptr()), 1.0,
              0.0, n, ctypes.c_void_p(out.data_ptr()), None)
    torch.cuda.synchronize()
    assert float(out.item()) == 0.0  # metrics.rs:19-21


def test_rollout_truth_matches_rollout_and_oracle(tiny):
    c = tiny.cfg
    nx, ny, n = c.nx, c.ny, c.nx * c.ny
    p = sw.Params(nx=nx, ny=ny, dx=c.dx, dt=c.dt, g=c.g, depth=c.depth, nu=c.nu)
    mean, std, depth = norm_of(tiny)
    st = sw.State.flat(p)
    st.add_drop(p, nx / 2.0, ny / 2.0, 0.05, 4.0)  # main.rs:63-78 "drop"
    x0 = O.encode(st.h, st.u, st.v, mean, std, depth)
    from jacobi2000_amd import checkpoint as ck

    meta = ck.meta_of(c.dx, c.dt, c.g, c.depth, c.nu)
    steps = 6
    solver = sw.ShallowWater(p)
    print("rollout truth: pt-tiny 16x16, 6 steps", flush=True)
    x, h, u, v, corr = tiny.rollout_truth(solver, x0, st.h, st.u, st.v, meta, steps)
    # the model half: the device rollout with the mass projection, bit for bit
    xr, _, _ = tiny.rollout(x0, meta, steps, project_mass=True, project_energy=False)
    assert (x.view(np.uint32) == xr.view(np.uint32)).all()
    # the truth half: the oracle's solver, bit for bit
    op = O.SwParams(nx, ny, p.dx, p.dt, p.g, p.depth, p.nu)
    oh, ou, ov = O.sw_steps(op, st.h, st.u, st.v, steps)
    assert (h.view(np.uint32) == oh.view(np.uint32)).all() and (u == ou).all() and (v == ov).all()
    # the trace: decode_h of each step's state against the solver's h - depth
    hh, uu, vv = st.h.copy(), st.u.copy(), st.v.copy()
    for t in range(1, steps + 1):
        xt, _, _ = tiny.rollout(x0, meta, t, project_mass=True, project_energy=False)
        hh, uu, vv = O.sw_steps(op, hh, uu, vv, 1)
        want = O.correlation(O.decode_h(xt, n, mean, std), hh - f32(depth))
        assert abs(corr[t - 1] - want) < 2e-6, (t, corr[t - 1], want)


def test_rollout_truth_rejects_a_solver_on_another_grid(tiny):
    c = tiny.cfg
    p = sw.Params(nx=c.nx + 1, ny=c.ny, dx=c.dx, dt=c.dt, g=c.g, depth=c.depth, nu=c.nu)
    n = c.nx * c.ny
    with pytest.raises(_lib.J2KError) as e:
        tiny.rollout_truth(sw.ShallowWater(p), np.zeros(3 * n, f32), np.ones(n, f32), np.zeros(n, f32),
                           np.zeros(n, f32), np.zeros(5, f32), 2)
    assert e.value.code == _lib.ERR_SHAPE
