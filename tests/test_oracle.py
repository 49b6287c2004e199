"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

- form P forward/rollout against the reference's PyTorch goldens (pt-tiny,
  crates/engine/tests/parity.rs:36-70 gates: 1e-5 one step, 1e-4 rollout);
- mask predicate: the truth table of crates/engine/src/mask.rs:38-57;
- projections: the known-answer tests of crates/engine/src/project.rs:64-102;
- GELU against double-precision erf;
- form C (no reference exists): an independent dense restatement with an
  explicit space-time mask matrix, translation equivariance of the periodic
  stencil, f32/f64 agreement, and projection invariants.
"""

import math

import numpy as np
import pytest

from jacobi2000_amd import checkpoint as ck
from oracle import oracle as O


@pytest.fixture(scope="module")
def golden(fixture_dir):
    return O.read_safetensors((fixture_dir / "reference.safetensors").read_bytes())


@pytest.fixture(scope="module")
def tiny(fixture_dir):
    return O.FormP.load(fixture_dir)


@pytest.mark.parametrize("f64", [False, True])
def test_stencil_one_step_matches_pytorch(tiny, golden, f64):
    out = tiny.forward(golden["state"], golden["meta"], f64)
    assert np.abs(out - golden["forward_stencil"].reshape(-1)).max() < 1e-5


@pytest.mark.parametrize("f64", [False, True])
def test_full_one_step_matches_pytorch(fixture_dir, golden, f64):
    m = O.FormP.load((fixture_dir / "model.safetensors").read_bytes(), (fixture_dir / "config-full.json").read_text())
    out = m.forward(golden["state"], golden["meta"], f64)
    assert np.abs(out - golden["forward_full"].reshape(-1)).max() < 1e-5


def test_fifty_step_rollout_matches_pytorch(tiny, golden):
    s = golden["state"].reshape(-1).copy()
    for i in range(1, 51):
        s = tiny.step(s, golden["meta"])
        if i in (1, 10, 50):
            assert np.abs(s - golden[f"rollout_{i}"].reshape(-1)).max() < 1e-4, i


def test_a_dropped_term_fails_the_golden(tiny, golden):
    """A plausible mistake planted in the oracle's inputs (the pos term
    zeroed) must break the 1e-5 gate, so the gate has teeth."""
    bad = O.FormP(tiny.cfg, dict(tiny.w))
    bad.w["pos"] = np.zeros_like(bad.w["pos"])
    out = bad.forward(golden["state"], golden["meta"])
    assert np.abs(out - golden["forward_stencil"].reshape(-1)).max() > 1e-5


def test_stencil_truth_table():
    # mask.rs:38-47 on a 16x16 token grid
    tx, q = 16, 5 * 16 + 5
    a = lambda k: O.mask_allowed(1, 1, 1, tx, q, k)
    assert a(q) and a(5 * tx + 6) and a(6 * tx + 6)
    assert not a(5 * tx + 7) and not a(9 * tx + 5)
    # mask.rs:49-56 dilation
    d = lambda k: O.mask_allowed(1, 1, 2, tx, q, k)
    assert d(5 * tx + 7) and not d(5 * tx + 6)
    assert O.mask_allowed(0, 0, 1, tx, 0, 255)


def test_mass_projection_restores_target_and_leaves_uv():
    # project.rs:64-72
    n = 4
    x = np.array([0.1, 0.2, 0.3, 0.4, 5, 5, 5, 5, -1, -1, -1, -1], np.float32)
    u = x[n:2 * n].copy()
    O.project_mass(x, 0.0, n)
    assert abs(O.mass(x, n)) < 1e-6
    assert np.array_equal(x[n:2 * n], u)


def test_energy_projection_scales_down_to_bound():
    # project.rs:74-89: potential 78.48 + kinetic 64 over a bound of 100
    p = O.phys(g=9.81)
    n = 16
    x = np.concatenate([np.ones(n), np.full(n, 2.0), np.full(n, 2.0)]).astype(np.float32)
    O.project_energy(x, 1.0, O.phys(g=9.81, r_max=100.0), n)
    e = O.energy(x, p, n)
    assert e <= 100.0 + 1e-3 and e > 99.0
    assert abs(O.mass(x, n) - 1.0) < 1e-6


def test_energy_projection_identity_within_bound():
    # project.rs:91-101
    n = 16
    x = np.concatenate([np.ones(n), np.full(n, 2.0), np.full(n, 2.0)]).astype(np.float32)
    before = x.copy()
    O.project_energy(x, 2.0, O.phys(g=9.81, r_max=100.0), n)
    assert np.array_equal(x, before)


def test_energy_closed_form():
    # E = sum 1/2 g h^2 + 1/2 h (u^2+v^2) in physical units (project.rs:9-25)
    p = O.phys(mean=(0.1, -0.05, 0.02), std=(0.3, 0.2, 0.2), depth=1.0, g=9.81)
    rng = np.random.default_rng(3)
    x = rng.normal(0, 0.2, 3 * 64).astype(np.float64)
    f = lambda v: float(np.float32(v))  # the physics constants are f32 in config.json's schema
    h = x[:64] * f(0.3) + (f(0.1) + 1.0)
    u = x[64:128] * f(0.2) + f(-0.05)
    v = x[128:] * f(0.2) + f(0.02)
    want = (0.5 * f(9.81) * h * h + 0.5 * h * (u * u + v * v)).sum()
    assert abs(O.energy(x, p, 64) - want) < 1e-9 * want


def test_gelu_matches_double_erf():
    for x in np.linspace(-6, 6, 241):
        want = 0.5 * x * (1 + math.erf(x / math.sqrt(2)))
        assert abs(O.gelu(float(x)) - want) <= 2e-7 * max(1.0, abs(want))


def test_safetensors_round_trip():
    # weights.rs:60-78 with a 2x2 tensor, and byte compatibility with the safetensors package
    t = {"w": np.array([[1, 2], [3, 4]], np.float32), "b": np.arange(3, dtype=np.float32)}
    blob = ck.write_safetensors(t)
    back = O.read_safetensors(blob)
    assert np.array_equal(back["w"], t["w"]) and np.array_equal(back["b"], t["b"])
    st = pytest.importorskip("safetensors.numpy")
    lib_back = st.load(blob)
    assert np.array_equal(lib_back["w"], t["w"])
    assert np.array_equal(O.read_safetensors(st.save(t))["b"], t["b"])


# ---------------------------------------------------------------- form C

def _dense_cell_forward(m: O.FormC, grid, known, meta):
    """Independent f64 restatement of the cell_st forward: the stencil is an
    explicit [M, M] mask built from periodic distances, attention is dense."""
    w = {k: v.astype(np.float64) for k, v in m.w.items()}
    _, T, ny, nx = grid.shape
    M = T * ny * nx
    t, y, x = np.meshgrid(np.arange(T), np.arange(ny), np.arange(nx), indexing="ij")
    t, y, x = t.reshape(-1), y.reshape(-1), x.reshape(-1)
    st = m.cfg["stencil"]
    dy = np.abs(y[:, None] - y[None, :])
    dx = np.abs(x[:, None] - x[None, :])
    dy = np.minimum(dy, ny - dy)
    dx = np.minimum(dx, nx - dx)
    allowed = (dy <= st["radius"]) & (dx <= st["radius"]) & (np.abs(t[:, None] - t[None, :]) <= st["time_radius"])
    feat = np.stack([grid[0].reshape(-1), grid[1].reshape(-1), grid[2].reshape(-1),
                     known.reshape(-1).astype(np.float64)], 1)
    X = feat @ w["embed.weight"] + w["embed.bias"] + (meta.astype(np.float64) @ w["meta.weight"] + w["meta.bias"])

    def ln(z, g, b):
        mu = z.mean(1, keepdims=True)
        var = ((z - mu) ** 2).mean(1, keepdims=True)
        return (z - mu) / np.sqrt(var + 1e-5) * g + b

    H, d = m.cfg["heads"], m.cfg["dim"]
    hd = d // H
    gelu = np.vectorize(lambda v: 0.5 * v * (1 + math.erf(v / math.sqrt(2))))
    for i in range(m.cfg["layers"]):
        p = lambda s: w[f"blocks.{i}.{s}"]
        h = ln(X, p("norm1.gamma"), p("norm1.beta"))
        q, k, v = (h @ p(f"{n}.weight") + p(f"{n}.bias") for n in "qkv")
        out = np.zeros_like(X)
        for hh in range(H):
            sl = slice(hh * hd, (hh + 1) * hd)
            s = q[:, sl] @ k[:, sl].T / math.sqrt(hd)
            s = np.where(allowed, s, -np.inf)
            s = np.exp(s - s.max(1, keepdims=True))
            out[:, sl] = (s / s.sum(1, keepdims=True)) @ v[:, sl]
        X = X + out @ p("proj.weight") + p("proj.bias")
        h = ln(X, p("norm2.gamma"), p("norm2.beta"))
        X = X + gelu(h @ p("ff1.weight") + p("ff1.bias")) @ p("ff2.weight") + p("ff2.bias")
    return ln(X, w["norm.gamma"], w["norm.beta"]) @ w["readout.weight"] + w["readout.bias"]


def _tiny_cell(seed=0, std=0.2, rt=1, rs=1):
    st, cj = ck.random_cell_st(dim=16, layers=2, heads=2, ff=32, seed=seed, std=std, radius=rs, time_radius=rt)
    return O.FormC.load(st, cj)


@pytest.mark.parametrize("shape", [(3, 4, 5), (4, 3, 3), (1, 6, 4)])
def test_cell_forward_matches_dense_restatement(shape):
    T, ny, nx = shape
    m = _tiny_cell()
    rng = np.random.default_rng(1)
    g = rng.normal(0, 0.3, (3, T, ny, nx)).astype(np.float32)
    k = (rng.random((T, ny, nx)) < 0.3).astype(np.uint8)
    meta = ck.meta_of(1.0, 0.09, 9.81, 1.0, 1e-3)
    y64 = m.forward(g, k, meta, f64=True)
    want = _dense_cell_forward(m, g.astype(np.float64), k, meta)
    assert np.abs(y64 - want).max() < 1e-10
    y32 = m.forward(g, k, meta)
    assert np.abs(y32 - want).max() < 1e-5


def test_cell_forward_wrong_wrap_would_fail():
    """The dense restatement distinguishes periodic from clipped space: a
    clipped stencil (dilation-free mask.rs rule) gives a different answer."""
    m = _tiny_cell()
    rng = np.random.default_rng(2)
    g = rng.normal(0, 0.3, (3, 2, 4, 4)).astype(np.float32)
    k = np.zeros((2, 4, 4), np.uint8)
    meta = ck.meta_of(1.0, 0.09, 9.81, 1.0, 1e-3)
    y = m.forward(g, k, meta, f64=True)
    m2 = _tiny_cell()
    m2.cfg = dict(m2.cfg, stencil={"radius": 1, "time_radius": 0})
    assert np.abs(m2.forward(g, k, meta, f64=True) - y).max() > 1e-6


def test_cell_forward_translation_equivariant():
    """No position term and a periodic stencil: shifting the grid cyclically
    in x and y shifts the output the same way."""
    m = _tiny_cell(seed=4)
    rng = np.random.default_rng(5)
    T, ny, nx = 2, 5, 6
    g = rng.normal(0, 0.3, (3, T, ny, nx)).astype(np.float32)
    k = (rng.random((T, ny, nx)) < 0.5).astype(np.uint8)
    meta = ck.meta_of(1.0, 0.09, 9.81, 1.0, 1e-3)
    y = m.forward(g, k, meta, f64=True).reshape(T, ny, nx, 3)
    gs = np.roll(g, (2, 3), axis=(2, 3))
    ks = np.roll(k, (2, 3), axis=(1, 2))
    ys = m.forward(gs, ks, meta, f64=True).reshape(T, ny, nx, 3)
    assert np.abs(np.roll(y, (2, 3), axis=(1, 2)) - ys).max() < 1e-12


def test_fill_blank_rules():
    T, ny, nx = 4, 2, 2
    g = np.arange(3 * T * ny * nx, dtype=np.float32).reshape(3, T, ny, nx)
    k = np.zeros((T, ny, nx), np.uint8)
    k[1] = 1
    k[3] = 1
    out = O.fill_blank(g, k)
    assert np.array_equal(out[:, 0], g[:, 1])  # nearest full slice
    assert np.array_equal(out[:, 2], g[:, 1])  # tie goes to the earlier slice
    k2 = np.zeros((T, ny, nx), np.uint8)
    k2[0, 0, 0] = k2[0, 1, 1] = 1
    out2 = O.fill_blank(g, k2)
    assert np.allclose(out2[:, 0, 0, 1], (g[:, 0, 0, 0] + g[:, 0, 1, 1]) / 2)  # slice mean of known cells
    assert np.allclose(out2[:, 2, 1, 0], (g[:, 0, 0, 0] + g[:, 0, 1, 1]) / 2)  # global mean for empty slices
    assert np.array_equal(out2[:, 0, 0, 0], g[:, 0, 0, 0])


def _rand_state(T, ny, nx, seed, uv=0.3):
    rng = np.random.default_rng(seed)
    g = np.zeros((3, T, ny, nx), np.float32)
    g[0] = rng.normal(0, 0.1, (T, ny, nx))
    g[1:] = rng.normal(0, uv, (2, T, ny, nx))
    return g


@pytest.mark.parametrize("f64", [False, True])
def test_cell_projection_invariants(f64):
    T, ny, nx = 3, 8, 8
    n = ny * nx
    p = O.phys(mean=(0.1, -0.05, 0.02), std=(0.3, 0.2, 0.2), depth=1.0, g=9.81, r_max=1.0)
    x0 = _rand_state(T, ny, nx, 0)
    mt, et = O.slice_targets(x0, p)
    k = np.zeros((T, ny, nx), np.uint8)
    k[:, ::3, ::2] = 1
    x = (x0 + _rand_state(T, ny, nx, 1, uv=0.5)).astype(np.float64 if f64 else np.float32)
    before = x.copy()
    O.project_c(x, k, p, mt, et)
    kb = k.astype(bool)
    assert np.array_equal(x[:, kb], before[:, kb])  # known cells untouched
    for s in range(T):
        assert abs(x[0, s].astype(np.float64).mean() - mt[s]) < 1e-6  # mass restored
    assert O.energy_violation(x.astype(np.float32), p, et) <= 1e-6


def test_cell_projection_identity_within_bound():
    T, ny, nx = 2, 6, 6
    p = O.phys(g=9.81, depth=1.0, r_max=1.0)
    x0 = _rand_state(T, ny, nx, 2)
    mt, et = O.slice_targets(x0, p)
    k = np.zeros((T, ny, nx), np.uint8)
    x = x0.copy()
    x[1:] *= 0.5  # less kinetic energy, same mass
    before = x.copy()
    O.project_c(x, k, p, mt, et)
    assert np.array_equal(x, before)


def test_cell_projection_reduces_to_reference_when_all_blank():
    """With every cell blank and enough potential budget, the cell_st energy
    step is project.rs:43-54 with the bound shrunk by (1 - 1e-7)."""
    T, ny, nx = 1, 8, 8
    n = ny * nx
    p = O.phys(g=9.81, depth=1.0, r_max=1.0)
    x0 = _rand_state(T, ny, nx, 3)
    mt, et = O.slice_targets(x0, p)
    x = x0.astype(np.float64).copy()
    x[1:] *= 1.5
    ref = x.reshape(3, n).copy().reshape(-1)
    O.project_c(x, np.zeros((T, ny, nx), np.uint8), p, mt, et)
    O.project_energy(ref, et[0] * (1 - 1e-7), p, n)
    assert np.abs(x.reshape(-1) - ref).max() < 1e-12


def test_cell_iterate_zero_readout_is_a_fixed_point():
    """readout = 0 means f = 0: x1 = P(x0) = x0 and the iteration stops after one step."""
    m = _tiny_cell()
    m.w = dict(m.w)
    m.w["readout.weight"] = np.zeros_like(m.w["readout.weight"])
    m.w["readout.bias"] = np.zeros_like(m.w["readout.bias"])
    g = ck.drops_grid(3, 6, 6, 2, 0)
    k = ck.mask_of("forward", 3, 6, 6)
    x0 = O.fill_blank(g, k)
    x, its, tr = m.iterate(x0, k, ck.meta_of(1.0, 0.09, 9.81, 1.0, 1e-3), 5, 1e-4)
    assert its == 1 and tr[0] == 0.0 and np.array_equal(x, x0)


def test_cell_iterate_f32_tracks_f64():
    st, cj = ck.random_cell_st(dim=32, layers=2, heads=2, ff=64, seed=0)
    m = O.FormC.load(st, cj)
    T, ny, nx = 4, 8, 8
    g = ck.drops_grid(T, ny, nx, 1, 0, centred=True)
    meta = ck.meta_of(1.0, 0.09, 9.81, 1.0, 1e-3)
    for kind in ("forward", "rewind", "reconstruction"):
        k = ck.mask_of(kind, T, ny, nx, seed=1, frac=0.1)
        x0 = O.fill_blank(g, k)
        x32, it32, _ = m.iterate(x0, k, meta, 6, 1e-4)
        x64, _, _ = m.iterate(x0, k, meta, it32, 0.0, f64=True)
        assert np.abs(x32 - x64).max() < 1e-5
        mt, et = O.slice_targets(x0, m.phys())
        assert O.energy_violation(x32, m.phys(), et) <= 1e-6
