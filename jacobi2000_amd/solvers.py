"""Python mirror of the reference `solvers` crate's shallow-water API over
libj2k's device solver (crates/solvers/src/shallow_water.rs):

    Params { nx, ny, dx, dt, g, depth, nu }, Params::default(), cfl_dt,
    with_cfl_dt, cells, cells_per_tick, horizon_ticks          (:16-80)
    State { h, u, v }, State::flat, State::mass, State::add_drop   (:82-112)
    ShallowWater::new(params), step(&mut State), steps(&mut State, n)
                                                                 (:126-244)

Every step runs on the GPU and is bit-identical to the reference's f32
arithmetic. `ShallowWater.trajectories` is dump-dataset's loop
(crates/engine-cli/src/dump_dataset.rs:168-193) over many states at once.
The scalar helpers of Params compute in f32, as the reference does.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, replace

import numpy as np

from . import _lib
from ._lib import call

_f = np.float32


@dataclass(frozen=True)
class Params:
    nx: int = 64
    ny: int = 64
    dx: float = 1.0
    dt: float | None = None  # None: Params::default()'s dt, as `..Default::default()`
    g: float = 9.81
    depth: float = 1.0
    nu: float = 0.0

    def __post_init__(self):
        if self.dt is None:  # default() is with_cfl_dt(0.4) of the default grid (:31-44)
            object.__setattr__(self, "dt", float(Params(dt=0.0).cfl_dt(0.4)))

    def cfl_dt(self, safety: float) -> float:
        """shallow_water.rs:47-52: safety * dx / (sqrt(g*depth) * SQRT_2), f32."""
        c = np.sqrt(_f(self.g) * _f(self.depth), dtype=_f)
        return float(_f(safety) * _f(self.dx) / (c * _f(np.sqrt(2.0))))

    def with_cfl_dt(self, safety: float) -> "Params":
        return replace(self, dt=self.cfl_dt(safety))

    def cells(self) -> int:
        return self.nx * self.ny

    def cells_per_tick(self) -> float:
        """shallow_water.rs:63-66."""
        return float(np.sqrt(_f(self.g) * _f(self.depth), dtype=_f) * _f(self.dt) / _f(self.dx))

    def horizon_ticks(self, cx: float, cy: float) -> float:
        """shallow_water.rs:68-79."""
        dist = min(_f(cx), _f(cy), _f(self.nx) - _f(cx), _f(self.ny) - _f(cy))
        return float(_f(dist) / _f(self.cells_per_tick()))

    def c(self) -> _lib.SwParams:
        return _lib.SwParams(self.nx, self.ny, self.dx, self.dt, self.g, self.depth, self.nu)


def drop_field(p: Params, cx: float, cy: float, amp: float, sigma: float) -> np.ndarray:
    """The bump State::add_drop adds (shallow_water.rs:102-111), [ny, nx] f32."""
    out = np.empty((p.ny, p.nx), dtype=np.float32)
    call("j2k_sw_drop_field", ctypes.byref(p.c()), cx, cy, amp, sigma, out.ctypes.data_as(ctypes.c_void_p))
    return out


@dataclass
class State:
    h: np.ndarray
    u: np.ndarray
    v: np.ndarray

    @staticmethod
    def flat(p: Params) -> "State":
        n = p.cells()
        return State(np.full(n, p.depth, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32))

    def mass(self) -> float:
        """Sum of h in f64, left to right (shallow_water.rs:96-98)."""
        return float(np.cumsum(self.h.astype(np.float64))[-1]) if self.h.size else 0.0

    def add_drop(self, p: Params, cx: float, cy: float, amp: float, sigma: float) -> None:
        self.h += drop_field(p, cx, cy, amp, sigma).reshape(-1)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class ShallowWater:
    """A solver for one grid on one GPU (ShallowWater::new)."""

    RESIDENT_MAX_CELLS = 6826

    def __init__(self, params: Params, device: int = 0):
        self.params = params
        self._h = ctypes.c_void_p()
        call("j2k_sw_create", ctypes.byref(params.c()), device, ctypes.byref(self._h))

    def close(self) -> None:
        if self._h:
            call("j2k_sw_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_kernel(self, kernel: str) -> None:
        """"auto", "resident" (whole run in LDS, grids up to 6,826 cells) or
        "tiled" (one launch per step through HBM); results are identical."""
        call("j2k_sw_set_kernel", self._h, {"auto": 0, "resident": 1, "tiled": 2}[kernel])

    def step(self, s: State) -> None:
        self.steps(s, 1)

    def steps(self, s: State, n: int) -> None:
        """In place on one State, or on a batch whose fields are [B, ny, nx]."""
        for a in (s.h, s.u, s.v):
            assert a.dtype == np.float32 and a.flags.c_contiguous
        cells = self.params.cells()
        batch = s.h.size // cells
        assert s.h.size == batch * cells == s.u.size == s.v.size
        call("j2k_sw_steps", self._h, _ptr(s.h), _ptr(s.u), _ptr(s.v), batch, n)

    def steps_dev(self, d_h: int, d_u: int, d_v: int, batch: int, n: int, stream=None) -> None:
        """Device pointers (ints), fields [batch, ny, nx], in place."""
        call("j2k_sw_steps_dev", self._h, ctypes.c_void_p(d_h), ctypes.c_void_p(d_u), ctypes.c_void_p(d_v), batch,
             n, ctypes.c_void_p(stream))

    def trajectories(self, s: State, ticks, drops=(), fields=None) -> list[np.ndarray]:
        """dump-dataset's loop over a batch: state b (fields [B, ny, nx])
        runs ticks[b] steps; drops are (state, tick, field) triples into
        fields [k, ny, nx] (see drop_field), applied after that tick's step.
        Returns one [3, ticks[b]+1, ny, nx] array per state (the
        traj-NNNNN.f32 layout); the states advance in place."""
        p = self.params
        n = p.cells()
        t = np.ascontiguousarray(ticks, dtype=np.int32).reshape(-1)
        B = t.size
        assert s.h.size == B * n
        sizes = [3 * (int(k) + 1) * n for k in t]
        traj = np.empty(sum(sizes), dtype=np.float32)
        d = np.ascontiguousarray(np.array(drops, dtype=np.int32).reshape(-1, 3))
        fl = (np.zeros((0, p.ny, p.nx), np.float32) if fields is None
              else np.ascontiguousarray(fields, dtype=np.float32).reshape(-1, p.ny, p.nx))
        call("j2k_sw_trajectories", self._h, _ptr(s.h), _ptr(s.u), _ptr(s.v), B, _ptr(t), _ptr(traj), len(d),
             _ptr(d), _ptr(fl), fl.shape[0])
        out, o = [], 0
        for k, sz in zip(t, sizes):
            out.append(traj[o:o + sz].reshape(3, int(k) + 1, p.ny, p.nx))
            o += sz
        return out
