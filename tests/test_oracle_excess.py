"""Pins the oracle's excess-energy band (oracle/excess_oracle.c, an f64
restatement of training/jacobi2000_pt/project.py:66-140) before it checks
the device projection.

- The reference's band tests (training/tests/test_project.py:33-66) with
  their inputs: the band is kept, mass is exact, members inside the band are
  untouched, the upper-only form with r_min = 0.
- Identities that a dropped term, a wrong sign or factor breaks:
  * dt/dx = 0: excess = energy(x) - g n hbar^2 / 2, against the pinned
    project.rs energy (the mean of eta is zero);
  * the cross term equals g/2 sum eta (h_new - h), h_new the pinned
    solver's continuity step from the same state (the face fluxes are the
    solver's own);
  * scaling (eta, u, v) by s gives exactly s^2 a + s^3 q.
"""

import numpy as np
import pytest

from oracle import oracle as o

DT = 0.0903
G = 9.81
MEAN, STD, DEPTH = (0.0008, 0.0004, 0.0005), (0.0063, 0.0138, 0.0135), 1.0  # test_project.py:20


def phys():
    return o.phys(mean=MEAN, std=STD, depth=DEPTH, g=G)


def wavy(b):
    """test_project.py:24-28: [b, 3, 8, 8]."""
    n = 3 * 8 * 8
    i = np.arange(b * n, dtype=np.float32)
    amp = 1.0 + (i // n)
    return (amp * (np.sin(i * np.float32(0.37)) + np.float32(0.5) * np.cos(i * np.float32(1.91)))).reshape(
        b, 3, 8, 8).astype(np.float64)


def mass(x):
    return x[0].mean()


def test_band_is_kept_and_mass_exact():
    x = wavy(6)
    p = phys()
    r_min, r_max = 0.9999, 1.0004
    for i, f in enumerate([2.0, 1.0002, 1.0, 0.9995, 0.5, 0.0]):
        e = o.excess_energy(x[i], 8, 8, p, DT)
        prev = e * f
        y, _ = o.project_excess_energy(x[i], 8, 8, p, DT, prev, r_min, r_max)
        e1 = o.excess_energy(y, 8, 8, p, DT)
        assert e1 >= prev * r_min - 1e-12 * abs(prev), i
        assert e1 <= prev * r_max + 1e-12, i
        assert abs(mass(y.reshape(3, 8, 8)) - mass(x[i])) < 1e-12


def test_upper_only_form_and_identity_inside():
    x = wavy(3)
    p = phys()
    e0 = o.excess_energy(x[0], 8, 8, p, DT)
    y, s = o.project_excess_energy(x[0], 8, 8, p, DT, 2.0 * e0, 0.0, 1.0)
    assert s == 1.0 and (y == x[0].reshape(-1)).all()
    e1 = o.excess_energy(x[1], 8, 8, p, DT)
    y, s = o.project_excess_energy(x[1], 8, 8, p, DT, 0.9 * e1, 0.0, 1.0)
    assert s < 1.0 and o.excess_energy(y, 8, 8, p, DT) <= 0.9 * e1
    for i in range(3):
        e = o.excess_energy(x[i], 8, 8, p, DT)
        y, s = o.project_excess_energy(x[i], 8, 8, p, DT, e, 0.9999, 1.0004)
        assert s == 1.0 and (y == x[i].reshape(-1)).all()


# A normalizer with large physical anomalies (eta ~ 0.1, |u| ~ 0.3) and a
# depth far from 1, so that every term (q, the hbar factors) is well above
# the tolerances; the identities hold for any normalizer.
BIG = dict(mean=(0.01, -0.02, 0.03), std=(0.1, 0.3, 0.25), depth=3.0, g=G)


def test_plain_excess_is_energy_less_the_mean_surface():
    rng = np.random.default_rng(0)
    nx, ny = 12, 10
    n = nx * ny
    p = o.phys(**BIG)
    x = rng.standard_normal((3, ny, nx))
    m = x[0].mean()
    a, q = o.excess_terms(x, nx, ny, p, 0.0)
    assert abs(q) > 1e-4 * abs(a)
    # the normalizer travels as f32 (as in config.json's reader)
    std0, mean0, depth, g = (float(np.float32(v)) for v in (BIG["std"][0], BIG["mean"][0], BIG["depth"], G))
    hbar = m * std0 + (mean0 + depth)
    # project.rs energy (pinned by its known-answer tests), on the same f64 state
    import ctypes
    xx = np.ascontiguousarray(x.reshape(-1), dtype=np.float64)
    e_tot = o.lib().orc_energy_f64(xx.ctypes.data_as(ctypes.c_void_p), ctypes.byref(p), n)
    want = e_tot - 0.5 * g * n * hbar * hbar
    got = o.excess_energy(x, nx, ny, p, 0.0)
    assert abs(got - want) < 1e-9 * abs(e_tot)
    assert abs(got) > 1e-6  # not a vanishing difference


def test_cross_term_is_the_solvers_continuity_step():
    rng = np.random.default_rng(1)
    nx, ny = 16, 14
    n = nx * ny
    p = o.phys(**BIG)
    depth = BIG["depth"]
    m32 = [float(np.float32(v)) for v in BIG["mean"]]
    s32 = [float(np.float32(v)) for v in BIG["std"]]
    d32 = float(np.float32(depth))
    f32 = np.float32
    h = (depth + 0.1 * rng.standard_normal(n)).astype(f32)
    u = (0.3 * rng.standard_normal(n)).astype(f32)
    v = (0.3 * rng.standard_normal(n)).astype(f32)
    x = np.concatenate([(h.astype(np.float64) - m32[0] - d32) / s32[0],
                        (u.astype(np.float64) - m32[1]) / s32[1],
                        (v.astype(np.float64) - m32[2]) / s32[2]])
    cross = o.excess_energy(x, nx, ny, p, DT) - o.excess_energy(x, nx, ny, p, 0.0)
    # one solver step with dx = 1, dt = dt/dx: h_new = h - dt/dx div(F)
    sp = o.sw_params(nx=nx, ny=ny, dx=1.0, dt=DT, g=G, depth=depth, nu=0.0)
    h1, _, _ = o.sw_steps(sp, h, u, v, 1)
    eta = h.astype(np.float64) - h.astype(np.float64).mean()
    want = 0.5 * G * np.dot(eta, h1.astype(np.float64) - h.astype(np.float64))
    assert abs(cross - want) < 2e-3 * abs(want)
    assert abs(want) > 1e-4


def test_scaling_law():
    rng = np.random.default_rng(2)
    nx, ny = 10, 9
    n = nx * ny
    p = o.phys(**BIG)
    M = [float(np.float32(v)) for v in BIG["mean"]]
    S = [float(np.float32(v)) for v in BIG["std"]]
    x = rng.standard_normal((3, n))
    a, q = o.excess_terms(x, nx, ny, p, DT)
    assert abs(q) > 1e-4 * abs(a)
    m = x[0].mean()
    for s in (0.3, 0.97, 1.6):
        y = x.copy()
        y[0] = (x[0] - m) * s + m
        y[1] = ((x[1] * S[1] + M[1]) * s - M[1]) / S[1]
        y[2] = ((x[2] * S[2] + M[2]) * s - M[2]) / S[2]
        e = o.excess_energy(y, nx, ny, p, DT)
        assert abs(e - (s * s * a + s ** 3 * q)) < 1e-10 * (abs(a) + abs(q))


def test_projection_lands_on_its_target():
    # eta follows |u|^2, so q is a sizable share of the excess and the
    # Newton steps on s^2 a + s^3 q = target have real work to do
    rng = np.random.default_rng(3)
    nx, ny = 12, 12
    n = nx * ny
    p = o.phys(**BIG)
    M = [float(np.float32(v)) for v in BIG["mean"]]
    S = [float(np.float32(v)) for v in BIG["std"]]
    u = 0.5 * rng.standard_normal(n)
    v = 0.5 * rng.standard_normal(n)
    s2 = u * u + v * v
    eta = 0.2 * (s2 - s2.mean())
    x = np.concatenate([(eta + 0.01) / S[0] - (M[0]) / S[0], (u - M[1]) / S[1], (v - M[2]) / S[2]])
    a, q = o.excess_terms(x, nx, ny, p, DT)
    assert q > 0.02 * a
    e = a + q
    for f, (r_min, r_max) in [(0.5, (0.9999, 1.0004)), (2.0, (0.9999, 1.0004)), (0.7, (0.0, 1.0))]:
        prev = e * f
        y, s = o.project_excess_energy(x, nx, ny, p, DT, prev, r_min, r_max)
        lower, upper = prev * r_min, prev * r_max
        target = min(max(e, lower * (1 + 1e-5)), upper * (1 - 1e-5))
        got = o.excess_energy(y, nx, ny, p, DT)
        assert abs(got - target) < 1e-9 * target, (f, got, target)
