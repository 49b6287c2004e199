"""Python mirror of the reference `engine` crate API over libj2k.

Reference interface (crates/engine/src/lib.rs:30-68, gpu.rs:199-437):

    Model::load(dir) / Model::from_bytes(safetensors, config_json)
    Model::forward(state, meta) -> increment      Model::step(state, meta)
    Model::to_gpu() -> GpuModel                   GpuModel::{forward, step, cfg}
    project::{mass, energy, project_mass, project_energy}

Here every model lives on the GPU (there is no CPU path in the product):
`Model` and `GpuModel` are the same device-backed object. Loader failures
raise `J2KError` with the status the reference would have panicked with
(BadConfig, MissingTensor, BadDType, ShapeMismatch); a missing device
raises J2KError(NoAdapter) as `GpuModel::new` returns `Error::NoAdapter`.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import J2KError, call

PRECISIONS = {"f32": _lib.PREC_F32, "fp32": _lib.PREC_F32, "bf16": _lib.PREC_BF16, "bf16x2": _lib.PREC_BF16X2,
              "fp8": _lib.PREC_FP8}


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = _lib.lib().j2k_device_count(ctypes.byref(n))
    return n.value if rc == _lib.OK else 0


class Model:
    """A checkpoint on one GPU. `precision` is "f32" (plain form), "bf16x2"
    (served form: bf16 weights on MFMA, f32 activations as bf16 pairs), "bf16"
    (bf16 activations too) or "fp8" (the block products on E4M3 codes,
    weights with per-tensor scales, activations with per-token scales);
    every form accumulates in f32."""

    def __init__(self, handle: int, precision: str):
        self._h = ctypes.c_void_p(handle)
        self.precision = precision
        c = _lib.Config()
        call("j2k_model_config", self._h, ctypes.byref(c))
        self.cfg = c

    # ---- construction (lib.rs:36-50, gpu.rs:212-218) ----
    @staticmethod
    def load(dir: str | Path, precision: str = "f32", device: int = 0) -> "Model":
        d = Path(dir)
        return Model.from_bytes((d / "model.safetensors").read_bytes(), (d / "config.json").read_text(), precision,
                                device)

    @staticmethod
    def from_bytes(safetensors: bytes, config_json: str, precision: str = "f32", device: int = 0) -> "Model":
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        buf = (ctypes.c_uint8 * len(safetensors)).from_buffer_copy(safetensors)
        h = ctypes.c_void_p()
        call("j2k_model_create", buf, len(safetensors), config_json.encode(), device, PRECISIONS[precision],
             ctypes.byref(h))
        return Model(h.value, precision)

    def to_gpu(self) -> "Model":
        """lib.rs:65-67. The model already lives on the GPU."""
        return self

    def close(self) -> None:
        if self._h:
            call("j2k_model_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    @property
    def form(self) -> str:
        return "cell_st" if self.cfg.form == _lib.FORM_CELL_ST else "patch2d"

    # ---- patch2d: the reference forward ----
    def _grid_size(self) -> int:
        return self.cfg.channels * self.cfg.ny * self.cfg.nx

    def forward(self, state, meta) -> np.ndarray:
        """Increment for one normalized state [channels, ny, nx] (cpu.rs:249-352)."""
        s, m = _f32(state).reshape(-1), _f32(meta).reshape(-1)
        self._check_p(s, m)
        out = np.empty_like(s)
        call("j2k_forward", self._h, _ptr(s), _ptr(m), _ptr(out))
        return out

    def step(self, state, meta) -> np.ndarray:
        """state + forward(state) (cpu.rs:84-88)."""
        s, m = _f32(state).reshape(-1), _f32(meta).reshape(-1)
        self._check_p(s, m)
        out = np.empty_like(s)
        call("j2k_step", self._h, _ptr(s), _ptr(m), _ptr(out))
        return out

    def _check_p(self, s: np.ndarray, m: np.ndarray) -> None:
        if s.size != self._grid_size():
            raise J2KError(_lib.ERR_SHAPE, f"state has {s.size} values, expected {self._grid_size()}")
        if m.size != self.cfg.meta_dim:
            raise J2KError(_lib.ERR_SHAPE, f"meta has {m.size} values, expected {self.cfg.meta_dim}")

    def rollout(self, state, meta, steps: int, project_mass: bool = True, project_energy: bool = False):
        """The rollout/eval loop on the device (engine-cli main.rs:117-121, eval.rs:144-156).
        Returns (final state, energy trace [steps+1], max relative energy violation)."""
        s, m = _f32(state).reshape(-1).copy(), _f32(meta).reshape(-1)
        self._check_p(s, m)
        e = np.zeros(steps + 1, dtype=np.float64)
        v = ctypes.c_double(0)
        call("j2k_rollout", self._h, _ptr(s), _ptr(m), steps, int(project_mass), int(project_energy), _ptr(e),
             ctypes.byref(v))
        return s, e, v.value

    def rollout_truth(self, solver, state, h, u, v, meta, steps: int):
        """The rollout CLI's loop (engine-cli main.rs:116-127) with its solver
        on the device: the solver (solvers.ShallowWater on this grid) steps
        the physical (h, u, v), the model steps the normalized state with the
        mass projection, and each step's correlation of decode_h(state) with
        h - depth is recorded. Returns (state, h, u, v, corr [steps])."""
        x = _f32(state).reshape(-1).copy()
        hh, uu, vv = (_f32(a).reshape(-1).copy() for a in (h, u, v))
        m = _f32(meta).reshape(-1)
        self._check_p(x, m)
        corr = np.zeros(max(steps, 0), dtype=np.float32)
        call("j2k_rollout_truth", self._h, solver._h, _ptr(x), _ptr(m), _ptr(hh), _ptr(uu), _ptr(vv), steps,
             _ptr(corr))
        return x, hh, uu, vv, corr

    def score(self, x, h, u, v, centers, max_lag: int = 8) -> list[dict]:
        """eval-metrics' scores (engine-cli eval.rs:158-172) of a batch of
        model states x [B, 3, ny, nx] (normalized) against solver states
        h, u, v [B, ny, nx] (physical), drop centres [B, 2], on the device."""
        xx = _f32(x).reshape(-1)
        hh, uu, vv = (_f32(a).reshape(-1) for a in (h, u, v))
        n = self.cfg.nx * self.cfg.ny
        B = hh.size // n
        c = _f32(centers).reshape(-1)
        if xx.size != 3 * B * n or uu.size != B * n or vv.size != B * n or c.size != 2 * B:
            raise J2KError(_lib.ERR_SHAPE, "score: x must be [B, 3, ny, nx], h/u/v [B, ny, nx], centers [B, 2]")
        out = (_lib.Score * B)()
        call("j2k_score", self._h, B, _ptr(xx), _ptr(hh), _ptr(uu), _ptr(vv), _ptr(c), max_lag, out)
        return [{f: getattr(o, f) for f, _ in _lib.Score._fields_} for o in out]

    def iterate_window(self, window, meta, iters: int, project_mass: bool = False,
                       project_energy: bool = False) -> np.ndarray:
        """Jacobi-over-time window [T, channels, ny, nx] (slice 0 = initial state)."""
        w = _f32(window).copy()
        T = w.shape[0]
        m = _f32(meta).reshape(-1)
        call("j2k_iterate_window", self._h, _ptr(w), _ptr(m), T, iters, int(project_mass), int(project_energy))
        return w

    def mass(self, x, n: int) -> float:
        xx = _f32(x).reshape(-1)
        out = ctypes.c_double()
        call("j2k_mass", self._h, _ptr(xx), n, ctypes.byref(out))
        return out.value

    def energy(self, x, n: int) -> float:
        xx = _f32(x).reshape(-1)
        out = ctypes.c_double()
        call("j2k_energy", self._h, _ptr(xx), n, ctypes.byref(out))
        return out.value

    def project_mass(self, x: np.ndarray, target: float, n: int) -> None:
        """In place, project.rs:33-38."""
        assert x.dtype == np.float32 and x.flags.c_contiguous
        call("j2k_project_mass", self._h, _ptr(x), n, float(target))

    def project_energy(self, x: np.ndarray, prev_energy: float, n: int) -> None:
        """In place, project.rs:43-54 (r_max and the normalizer from config.json)."""
        assert x.dtype == np.float32 and x.flags.c_contiguous
        call("j2k_project_energy", self._h, _ptr(x), n, float(prev_energy))

    # ---- cell_st: per-cell space-time fixed point ----
    def iterate(self, grid, known, meta, max_iter: int = 0, tol: float = -1.0):
        """grid [3, T, ny, nx] (blank cells filled), known [T, ny, nx] ->
        (final grid, iterations, last max|dx|, energy violation)."""
        g = _f32(grid).copy()
        _, T, ny, nx = g.shape
        k = np.ascontiguousarray(known, dtype=np.uint8)
        m = _f32(meta).reshape(-1)
        it, r, v = ctypes.c_int(), ctypes.c_float(), ctypes.c_double()
        call("j2k_iterate", self._h, _ptr(g), _ptr(k), _ptr(m), T, ny, nx, max_iter, tol, ctypes.byref(it),
             ctypes.byref(r), ctypes.byref(v))
        return g, it.value, r.value, v.value

    def cell_forward_dev(self, d_grid, d_known, d_meta, T: int, ny: int, nx: int, d_y, stream=None) -> None:
        """y = f(x) on device pointers (ints), y [T*ny*nx, 3]."""
        call("j2k_cell_forward_dev", self._h, ctypes.c_void_p(d_grid), ctypes.c_void_p(d_known),
             ctypes.c_void_p(d_meta), T, ny, nx, ctypes.c_void_p(d_y), ctypes.c_void_p(stream))

    def iterate_dev(self, d_grid, d_known, d_meta, T: int, ny: int, nx: int, max_iter: int = 0, tol: float = -1.0,
                    stream=None):
        it, r, v = ctypes.c_int(), ctypes.c_float(), ctypes.c_double()
        call("j2k_iterate_dev", self._h, ctypes.c_void_p(d_grid), ctypes.c_void_p(d_known), ctypes.c_void_p(d_meta),
             T, ny, nx, max_iter, tol, ctypes.byref(it), ctypes.byref(r), ctypes.byref(v), ctypes.c_void_p(stream))
        return it.value, r.value, v.value

    def autotune(self, max_tokens: int = 0) -> dict:
        """Tune launch shapes up to `max_tokens` packed tokens (0: only report);
        returns the tuning table (see j2k_model_autotune)."""
        import json

        buf = ctypes.create_string_buffer(1 << 16)
        call("j2k_model_autotune", self._h, int(max_tokens), buf, len(buf))
        return json.loads(buf.value.decode())


GpuModel = Model


def fill_blank(grid, known) -> np.ndarray:
    """Blank-cell fill used before iterating (host, see j2k_fill_blank)."""
    g = _f32(grid)
    _, T, ny, nx = g.shape
    k = np.ascontiguousarray(known, dtype=np.uint8)
    out = np.empty_like(g)
    call("j2k_fill_blank", _ptr(g), _ptr(k), T, ny, nx, _ptr(out))
    return out
