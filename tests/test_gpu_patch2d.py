"""GPU parity of the patch2d form (the reference forward) through the C ABI.

Gates follow crates/engine/tests/parity.rs: one step against PyTorch's
goldens, the 50-step rollout, the full mask, and a dilated r=2 stencil
against the CPU forward. The HIP path is held to the CPU gates (1e-5 one
step, 1e-4 rollout) rather than the looser GPU ones (1e-4 / 1e-3).
"""

import numpy as np
import pytest

import jacobi2000_amd as j
from jacobi2000_amd import checkpoint as ck
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def golden(fixture_dir):
    return O.read_safetensors((fixture_dir / "reference.safetensors").read_bytes())


@pytest.fixture(scope="module")
def tiny(fixture_dir):
    return j.Model.load(fixture_dir)


def maxabs(a, b):
    return float(np.abs(np.asarray(a, np.float64).reshape(-1) - np.asarray(b, np.float64).reshape(-1)).max())


def test_stencil_one_step_matches_pytorch(tiny, golden):
    d = maxabs(tiny.forward(golden["state"], golden["meta"]), golden["forward_stencil"])
    print(f"stencil 1-step max_abs={d:.3e}")
    assert d < 1e-5


def test_full_mask_one_step_matches_pytorch(fixture_dir, golden):
    m = j.Model.from_bytes((fixture_dir / "model.safetensors").read_bytes(), (fixture_dir / "config-full.json").read_text())
    d = maxabs(m.forward(golden["state"], golden["meta"]), golden["forward_full"])
    print(f"full 1-step max_abs={d:.3e}")
    assert d < 1e-5


def test_step_is_state_plus_forward(tiny, golden):
    s = golden["state"].reshape(-1)
    assert maxabs(tiny.step(s, golden["meta"]), s + tiny.forward(s, golden["meta"])) < 1e-6


def test_fifty_step_rollout_matches_pytorch(tiny, golden):
    s = golden["state"].reshape(-1).copy()
    for i in range(1, 51):
        s = tiny.step(s, golden["meta"])
        if i in (1, 10, 50):
            d = maxabs(s, golden[f"rollout_{i}"])
            print(f"rollout {i}: max_abs={d:.3e}")
            assert d < 1e-4


def test_device_rollout_matches_pytorch(tiny, golden):
    s, e, _ = tiny.rollout(golden["state"], golden["meta"], 50, project_mass=False, project_energy=False)
    assert maxabs(s, golden["rollout_50"]) < 1e-4
    assert e.shape == (51,) and np.all(e > 0)


def test_dilated_stencil_matches_cpu(fixture_dir, golden):
    import json
    cfg = json.loads((fixture_dir / "config.json").read_text())
    cfg["mask"] = {"Stencil": {"radius": 2, "dilation": 2}}
    st = (fixture_dir / "model.safetensors").read_bytes()
    gpu = j.Model.from_bytes(st, json.dumps(cfg))
    cpu = O.FormP.load(st, json.dumps(cfg))
    d = maxabs(gpu.forward(golden["state"], golden["meta"]), cpu.forward(golden["state"], golden["meta"]))
    print(f"r2 d2 max_abs={d:.3e}")
    assert d < 1e-5


def test_window_iteration_reaches_the_rollout(tiny, golden):
    """Jacobi over time: after k iterations slice k equals the k-step rollout."""
    T = 11
    w = np.zeros((T, 3, 16, 16), np.float32)
    w[:] = golden["state"]
    meta = golden["meta"]
    seq = [golden["state"].reshape(-1)]
    for _ in range(T - 1):
        seq.append(tiny.step(seq[-1], meta))
    for k in (1, 4, 10):
        out = tiny.iterate_window(w, meta, k)
        for t in range(k + 1):
            assert maxabs(out[t], seq[t]) < 1e-6, (k, t)
    assert maxabs(tiny.iterate_window(w, meta, 10)[10], golden["rollout_10"]) < 1e-4


def test_projected_rollout_matches_oracle(fixture_dir, tiny, golden):
    """main.rs:117-121 + eval.rs:144-156 order (mass, then energy), 20 steps."""
    cpu = O.FormP.load(fixture_dir)
    p = cpu.phys()
    n = 256
    x = golden["state"].reshape(-1).copy()
    mass0 = O.mass(x, n)
    e_prev = O.energy(x, p, n)
    viol = 0.0
    for _ in range(20):
        nx_ = cpu.step(x, golden["meta"])
        O.project_mass(nx_, mass0, n)
        O.project_energy(nx_, e_prev, p, n)
        e = O.energy(nx_, p, n)
        viol = max(viol, max(0.0, (e - e_prev * p.r_max) / e_prev))
        x, e_prev = nx_, e
    s, etrace, gviol = tiny.rollout(golden["state"], golden["meta"], 20, True, True)
    d = maxabs(s, x)
    print(f"projected 20-step max_abs={d:.3e} violation gpu={gviol:.3e} cpu={viol:.3e}")
    assert d < 1e-4
    # pt-tiny's r_max is 0.999: where potential energy alone exceeds the bound
    # the reference projection cannot meet it either; the device must report
    # the same violation as the CPU loop.
    assert abs(gviol - viol) <= 1e-6


def test_device_projections_match_oracle(tiny):
    rng = np.random.default_rng(0)
    n = 256
    x = rng.normal(0, 0.3, 3 * n).astype(np.float32)
    cfgp = O.phys(tuple(tiny.cfg.norm_mean), tuple(tiny.cfg.norm_std), tiny.cfg.norm_depth, tiny.cfg.g,
                  tiny.cfg.energy_r_max)
    assert abs(tiny.mass(x, n) - O.mass(x.astype(np.float64), n)) < 1e-12
    e64 = O.energy(x.astype(np.float64), cfgp, n)
    assert abs(tiny.energy(x, n) - e64) < 1e-6 * e64
    a, b = x.copy(), x.copy()
    tiny.project_mass(a, 0.25, n)
    O.project_mass(b, 0.25, n)
    assert maxabs(a, b) < 1e-6
    a, b = x.copy(), x.copy()
    tiny.project_energy(a, e64 * 0.9, n)
    O.project_energy(b, e64 * 0.9, cfgp, n)
    assert maxabs(a, b) < 1e-5


@pytest.fixture(scope="module")
def model20():
    st, cj = ck.random_patch2d(seed=0)
    return st, cj


def test_model20_forward_matches_oracle(model20):
    """The reference model20 shape (64x64, p4, d128, 4 layers, stencil r1)."""
    st, cj = model20
    gpu = j.Model.from_bytes(st, cj)
    cpu = O.FormP.load(st, cj)
    rng = np.random.default_rng(7)
    state = rng.normal(0, 0.1, (3, 64, 64)).astype(np.float32)
    meta = ck.meta_of(1.0, 0.09, 9.81, 1.0, 1e-3)
    out = gpu.forward(state, meta).reshape(3, -1)
    ref = cpu.forward(state, meta).reshape(3, -1)
    per_field = np.abs(out - ref).max(1)
    print("model20 max_abs per field", per_field)
    assert per_field.max() < 1e-4


def test_model20_bf16_error_vs_f64(model20):
    st, cj = model20
    gpu = j.Model.from_bytes(st, cj, precision="bf16")
    cpu = O.FormP.load(st, cj)
    rng = np.random.default_rng(8)
    state = rng.normal(0, 0.1, (3, 64, 64)).astype(np.float32)
    meta = ck.meta_of(1.0, 0.09, 9.81, 1.0, 1e-3)
    out = gpu.forward(state, meta).astype(np.float64)
    ref = cpu.forward(state, meta, f64=True)
    rel = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    print(f"model20 bf16 rel L2 vs f64 = {rel:.3e}")
    assert rel < 2e-2


def test_wrong_state_size_is_shape_error(tiny, golden):
    with pytest.raises(j.J2KError) as e:
        tiny.forward(np.zeros(10, np.float32), golden["meta"])
    assert e.value.kind == "ShapeMismatch"
