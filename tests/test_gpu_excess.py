"""The excess-energy band on the device (j2k_excess_energy*,
j2k_project_excess_energy*) against the oracle's f64 restatement of
training/jacobi2000_pt/project.py:66-140, and the reference's own band
tests (training/tests/test_project.py:33-66) run on the device's output.

Tolerances: the device computes every per-cell value in f32 (as the
reference's tensor ops) and sums in f64, the oracle everything in f64, so
the excess agrees to f32 rounding of the cell values (1e-6 relative); the
projected states to that rounding carried through s (1e-6 relative); the
band is met with the reference's 1e-5 margin to spare."""

import ctypes
import json

import numpy as np
import pytest

import jacobi2000_amd as j
from jacobi2000_amd import _lib
from jacobi2000_amd import checkpoint as ck
from oracle import oracle as O

pytestmark = pytest.mark.gpu
f32 = np.float32
DT = 0.0903
G = 9.81
D20 = dict(mean=(0.0008, 0.0004, 0.0005), std=(0.0063, 0.0138, 0.0135), depth=1.0)  # test_project.py:20
BIG = dict(mean=(0.01, -0.02, 0.03), std=(0.1, 0.3, 0.25), depth=3.0)


def small_model(norm, excess=None):
    st, cj = ck.random_patch2d(nx=8, ny=8, patch=4, dim=32, layers=1, heads=4, ff=64, head_dim=8, seed=0)
    c = json.loads(cj)
    c["norm"] = {"mean": list(norm["mean"]), "std": list(norm["std"]), "depth": norm["depth"]}
    c["g"] = G
    if excess is not None:
        c["excess_r"], c["excess_r_min"] = excess
    return j.Model.from_bytes(st, json.dumps(c))


def phys(norm):
    return O.phys(mean=norm["mean"], std=norm["std"], depth=norm["depth"], g=G)


def wavy(b):
    """test_project.py:24-28."""
    n = 3 * 8 * 8
    i = np.arange(b * n, dtype=f32)
    amp = 1.0 + (i // n)
    return (amp * (np.sin(i * f32(0.37)) + f32(0.5) * np.cos(i * f32(1.91)))).reshape(b, 3, 8, 8).astype(f32)


def excess(m, x, dt_dx):
    x = np.ascontiguousarray(x, f32)
    B = x.shape[0]
    out = np.zeros(B, np.float64)
    _lib.call("j2k_excess_energy", m.handle, B, x.ctypes.data_as(ctypes.c_void_p), dt_dx,
              out.ctypes.data_as(ctypes.c_void_p))
    return out


def project(m, x, prev, dt_dx, r_min, r_max):
    y = np.ascontiguousarray(x, f32).copy()
    p = np.ascontiguousarray(prev, np.float64)
    _lib.call("j2k_project_excess_energy", m.handle, y.shape[0], y.ctypes.data_as(ctypes.c_void_p),
              p.ctypes.data_as(ctypes.c_void_p), dt_dx, r_min, r_max)
    return y


def test_config_carries_the_exported_band():
    m = small_model(D20, excess=(1.0004, 0.9999))
    assert m.cfg.excess_r == f32(1.0004) and m.cfg.excess_r_min == f32(0.9999)
    m2 = small_model(D20)
    assert np.isnan(m2.cfg.excess_r) and np.isnan(m2.cfg.excess_r_min)


@pytest.mark.parametrize("norm", [D20, BIG], ids=["d20", "big"])
@pytest.mark.parametrize("dt_dx", [DT, 0.0])
def test_excess_matches_the_oracle(norm, dt_dx):
    m = small_model(norm)
    rng = np.random.default_rng(4)
    x = np.concatenate([wavy(4), rng.standard_normal((4, 3, 8, 8)).astype(f32)])
    got = excess(m, x, dt_dx)
    p = phys(norm)
    for b in range(x.shape[0]):
        want = O.excess_energy(x[b].astype(np.float64), 8, 8, p, dt_dx)
        assert abs(got[b] - want) < 1e-6 * abs(want), (b, got[b], want)


def test_reference_band_tests_on_the_device():
    m = small_model(D20)
    p = phys(D20)
    x = wavy(6)
    e = excess(m, x, DT)
    r_min, r_max = 0.9999, 1.0004
    prev = e * np.array([2.0, 1.0002, 1.0, 0.9995, 0.5, 0.0])
    y = project(m, x, prev, DT, r_min, r_max)
    for i in range(6):
        e1 = O.excess_energy(y[i].astype(np.float64), 8, 8, p, DT)
        assert e1 >= prev[i] * r_min - 1e-12, i
        assert e1 <= prev[i] * r_max + 1e-12, i
        # mass: the mean of normalized h, to f32 rounding of the rescale
        assert abs(float(y[i, 0].astype(np.float64).mean() - x[i, 0].astype(np.float64).mean())) < 1e-6, i
    # identity inside the band (test_project.py:55-59): bit for bit
    z = project(m, wavy(3), excess(m, wavy(3), DT), DT, r_min, r_max)
    assert (z.view(np.uint32) == wavy(3).view(np.uint32)).all()
    # upper-only form (test_project.py:46-53)
    x2 = wavy(2)
    e2 = excess(m, x2, DT)
    y2 = project(m, x2, e2 * np.array([2.0, 0.9]), DT, 0.0, 1.0)
    assert (y2[0].view(np.uint32) == x2[0].view(np.uint32)).all()
    assert O.excess_energy(y2[1].astype(np.float64), 8, 8, p, DT) <= 0.9 * e2[1]


def test_projection_matches_the_oracle_with_large_anomalies():
    m = small_model(BIG)
    p = phys(BIG)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((5, 3, 8, 8)).astype(f32)
    e = excess(m, x, DT)
    prev = e * np.array([0.5, 0.8, 1.0, 1.3, 2.0])
    y = project(m, x, prev, DT, 0.9999, 1.0004)
    for b in range(5):
        want, s = O.project_excess_energy(x[b].astype(np.float64), 8, 8, p, DT, prev[b], 0.9999, 1.0004)
        got = y[b].reshape(-1).astype(np.float64)
        assert np.abs(got - want).max() < 1e-6 * (1.0 + np.abs(want).max()), b
        if s == 1.0:
            assert (y[b].view(np.uint32) == x[b].view(np.uint32)).all()


def test_device_pointer_forms():
    import torch

    m = small_model(BIG)
    x = torch.from_numpy(wavy(3)).cuda()
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    _lib.call("j2k_excess_energy_dev", m.handle, 3, ctypes.c_void_p(x.data_ptr()), DT, ctypes.c_void_p(out.data_ptr()),
              None)
    torch.cuda.synchronize()
    assert np.allclose(out.cpu().numpy(), excess(m, wavy(3), DT), rtol=0, atol=0)
    prev = out * torch.tensor([0.5, 1.0, 2.0], dtype=torch.float64, device="cuda")
    _lib.call("j2k_project_excess_energy_dev", m.handle, 3, ctypes.c_void_p(x.data_ptr()),
              ctypes.c_void_p(prev.data_ptr()), DT, 0.9999, 1.0004, None)
    torch.cuda.synchronize()
    want = project(m, wavy(3), prev.cpu().numpy(), DT, 0.9999, 1.0004)
    assert (x.cpu().numpy().view(np.uint32) == want.view(np.uint32)).all()


def test_bad_band_is_invalid():
    m = small_model(D20)
    with pytest.raises(_lib.J2KError) as e:
        project(m, wavy(1), np.ones(1), DT, 1.1, 1.0)
    assert e.value.code == _lib.ERR_INVALID
