"""Pins the oracle's field scores (oracle/metrics_oracle.c) before they check
the device scoring: the reference's metric tests (crates/engine/src/
metrics.rs:115-137) and closed forms that a wrong bin, lag sign, refinement
or normalisation breaks."""

import numpy as np
import pytest

from oracle import oracle as o

f32 = np.float32
A = np.array([0.1, 0.5, -0.3, 0.9, -0.7], f32)


def test_reference_metric_tests():
    assert abs(o.correlation(A, A) - 1.0) < 1e-6  # metrics.rs:120-124
    assert o.correlation(np.ones(5, f32), A) == 0.0  # metrics.rs:126-131
    assert o.bias(A, A) == 0.0  # metrics.rs:133-137


def test_peak_mean_std_closed_forms():
    assert o.peak(A) == f32(0.9)
    assert o.peak(-A) == f32(0.9)
    x = np.arange(1, 101, dtype=f32)
    assert o.mean(x) == 50.5
    assert abs(o.std_dev(x) - np.sqrt((100 ** 2 - 1) / 12.0)) < 1e-4  # population std of 1..N


def test_bias_of_a_shifted_field():
    rng = np.random.default_rng(0)
    s = rng.standard_normal(4096).astype(f32) * f32(0.02)
    for c in (0.001, -0.004, 0.02):
        got = o.bias(s + f32(c), s)
        assert abs(got - c / np.std(s.astype(np.float64))) < 1e-4 * max(1.0, abs(got))


def _bins(nx, ny, cx, cy):
    yy, xx = np.mgrid[0:ny, 0:nx]
    dx = xx.astype(f32) - f32(cx)
    dy = yy.astype(f32) - f32(cy)
    r = np.sqrt(dx * dx + dy * dy, dtype=f32)
    return (r + f32(0.5)).astype(np.int64)


@pytest.mark.parametrize("nx,ny,cx,cy", [(32, 32, 16.0, 16.0), (40, 24, 13.3, 9.7), (16, 20, 3.2, 10.0)])
def test_radial_profile_of_a_radial_field(nx, ny, cx, cy):
    # a field constant on each bin: the profile is that constant's |value|
    bins = _bins(nx, ny, cx, cy)
    vals = np.cos(np.arange(bins.max() + 1) * 0.7).astype(f32) * f32(0.05)
    f = vals[bins].astype(f32)
    prof = o.radial_profile(f, nx, ny, cx, cy)
    rmax = int(max(np.floor(min(cx, cy, nx - cx, ny - cy)), 1.0))
    assert prof.size == rmax + 1
    for b in range(rmax + 1):
        if (bins == b).any():
            assert prof[b] == pytest.approx(abs(float(vals[b])), rel=1e-12, abs=1e-15)
        else:
            assert prof[b] == 0.0


def test_radial_profile_brute_force_average():
    rng = np.random.default_rng(1)
    nx, ny, cx, cy = 30, 26, 14.6, 12.2
    f = rng.standard_normal(nx * ny).astype(f32)
    bins = _bins(nx, ny, cx, cy).reshape(-1)
    prof = o.radial_profile(f, nx, ny, cx, cy)
    for b in range(prof.size):
        sel = np.abs(f[bins == b]).astype(np.float64)
        want = sel.mean() if sel.size else 0.0
        assert prof[b] == pytest.approx(want, rel=1e-12)


def _bump(n, c, w=2.5):
    i = np.arange(n, dtype=np.float64)
    return np.exp(-((i - c) / w) ** 2)


@pytest.mark.parametrize("k", [-5, -2, 0, 1, 3, 6])
def test_xcorr_lag_recovers_integer_shifts(k):
    n = 40
    solver = _bump(n, 18.0)
    model = _bump(n, 18.0 + k)  # the model's front sits k cells further out
    # the parabola through lags k-1, k, k+1 is nearly symmetric (the
    # overlap windows differ by a cell), so the refined lag stays within 0.01
    assert abs(o.xcorr_lag(model, solver, 8) - k) < 1e-2


def test_xcorr_lag_refines_between_cells_and_clips_at_the_range():
    n = 40
    solver = _bump(n, 15.0)
    lag = o.xcorr_lag(_bump(n, 17.4), solver, 8)
    assert 2.0 < lag < 3.0 and abs(lag - 2.4) < 0.15
    # a best lag at the edge of the range is not refined
    assert o.xcorr_lag(_bump(n, 23.0), solver, 8) == 8.0
    assert o.xcorr_lag(_bump(n, 7.0), solver, 8) == -8.0


def test_encode_decode_round_trip():
    rng = np.random.default_rng(2)
    n = 256
    h = (1.0 + 0.05 * rng.standard_normal(n)).astype(f32)
    u = (0.01 * rng.standard_normal(n)).astype(f32)
    v = (0.01 * rng.standard_normal(n)).astype(f32)
    mean, std, depth = [0.1, -0.05, 0.02], [0.3, 0.2, 0.2], 1.0
    x = o.encode(h, u, v, mean, std, depth)
    assert np.abs(x[n:2 * n] * f32(0.2) + f32(-0.05) - u).max() < 1e-8
    hd = o.decode_h(x, n, mean, std)
    assert np.abs(hd - (h - f32(depth))).max() < 1e-7
