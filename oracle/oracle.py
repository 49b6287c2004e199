"""ctypes front of the CPU oracle (libj2k_oracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module as the checker. The product
(jacobi2000_amd) never imports it.

Checkpoints are read here with a small safetensors parser of its own
(crates/engine/src/weights.rs:29-48 describes the layout: 8-byte LE header
length, JSON header, raw LE buffer) so the oracle does not depend on the
product's loader.
"""

from __future__ import annotations

import ctypes
import json
import os
import struct
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libj2k_oracle.so"

BLOCK = ["norm1.gamma", "norm1.beta", "q.weight", "q.bias", "k.weight", "k.bias", "v.weight", "v.bias",
         "proj.weight", "proj.bias", "norm2.gamma", "norm2.beta", "ff1.weight", "ff1.bias", "ff2.weight", "ff2.bias"]


def p_names(layers: int) -> list[str]:
    """Form P weight order passed to orc_forward_p_*."""
    names = ["embed.weight", "embed.bias", "pos", "meta.weight", "meta.bias"]
    for i in range(layers):
        names += [f"blocks.{i}.{b}" for b in BLOCK]
    names += ["norm.gamma", "norm.beta", "head.token_mix.weight", "head.token_mix.bias", "head.to_cells.weight",
              "head.to_cells.bias", "head.conv1.weight", "head.conv1.bias", "head.conv2.weight", "head.conv2.bias"]
    return names


def c_names(layers: int) -> list[str]:
    """Form C weight order passed to orc_forward_c_* / orc_iterate_c_*."""
    names = ["embed.weight", "embed.bias", "meta.weight", "meta.bias"]
    for i in range(layers):
        names += [f"blocks.{i}.{b}" for b in BLOCK]
    names += ["norm.gamma", "norm.beta", "readout.weight", "readout.bias"]
    return names


def read_safetensors(data: bytes) -> dict[str, np.ndarray]:
    n = struct.unpack("<Q", data[:8])[0]
    header = json.loads(data[8:8 + n])
    out = {}
    for name, v in header.items():
        if name == "__metadata__":
            continue
        assert v["dtype"] == "F32", name
        a, b = v["data_offsets"]
        out[name] = np.frombuffer(data[8 + n + a:8 + n + b], dtype="<f4").reshape(v["shape"]).copy()
    return out


def write_safetensors(tensors: dict[str, np.ndarray]) -> bytes:
    header, off, blobs = {}, 0, []
    for name in sorted(tensors):
        a = np.ascontiguousarray(tensors[name], dtype="<f4")
        header[name] = {"dtype": "F32", "shape": list(a.shape), "data_offsets": [off, off + a.nbytes]}
        off += a.nbytes
        blobs.append(a.tobytes())
    h = json.dumps(header, separators=(",", ":")).encode()
    h += b" " * ((8 - len(h) % 8) % 8)
    return struct.pack("<Q", len(h)) + h + b"".join(blobs)


class CfgP(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int) for f in ("nx", "ny", "patch", "channels", "dim", "layers", "heads", "head_dim",
                                            "ff_inner", "meta_dim", "mask_kind", "radius", "dilation")]


class CfgC(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int) for f in ("nx", "ny", "T", "channels", "dim", "layers", "heads", "ff_inner",
                                            "meta_dim", "rs", "rt")]


class Phys(ctypes.Structure):
    _fields_ = [("mean", ctypes.c_float * 3), ("std", ctypes.c_float * 3), ("depth", ctypes.c_float),
                ("g", ctypes.c_float), ("r_max", ctypes.c_float)]


def phys(mean=(0.0, 0.0, 0.0), std=(1.0, 1.0, 1.0), depth=0.0, g=9.81, r_max=1.0) -> Phys:
    p = Phys()
    for i in range(3):
        p.mean[i] = mean[i]
        p.std[i] = std[i]
    p.depth, p.g, p.r_max = depth, g, r_max
    return p


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        _lib = ctypes.CDLL(str(LIB_PATH))
        _lib.orc_mass_f32.restype = ctypes.c_float
        _lib.orc_energy_f32.restype = ctypes.c_float
        _lib.orc_mass_f64.restype = ctypes.c_double
        _lib.orc_energy_f64.restype = ctypes.c_double
        _lib.orc_gelu_f32.restype = ctypes.c_float
        _lib.orc_gelu_f32.argtypes = [ctypes.c_float]
        _lib.orc_project_mass_f32.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_int]
        _lib.orc_project_energy_f32.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_int]
        _lib.orc_project_mass_f64.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        _lib.orc_project_energy_f64.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_int]
        _lib.orc_iterate_c_f32.argtypes = [ctypes.c_void_p] * 6 + [ctypes.c_int, ctypes.c_float, ctypes.c_void_p]
        _lib.orc_iterate_c_f64.argtypes = [ctypes.c_void_p] * 6 + [ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
        _lib.orc_iterate_c_states_f32.argtypes = ([ctypes.c_void_p] * 6 +
                                                  [ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p])
        _lib.orc_iterate_c_states_f64.argtypes = ([ctypes.c_void_p] * 6 +
                                                  [ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p])
        _lib.orc_energy_violation_c.restype = ctypes.c_double
        _lib.orc_energy_violation_c.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_void_p]
    return _lib


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _wptrs(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


@dataclass
class FormP:
    """Form P model: config.json dict + tensors (crates/engine/src/config.rs:70-100)."""
    cfg: dict
    w: dict

    @staticmethod
    def load(dir_or_bytes, config_json: str | None = None) -> "FormP":
        if config_json is None:
            d = Path(dir_or_bytes)
            return FormP(json.loads((d / "config.json").read_text()), read_safetensors((d / "model.safetensors").read_bytes()))
        return FormP(json.loads(config_json), read_safetensors(dir_or_bytes))

    def c_cfg(self) -> CfgP:
        j = self.cfg
        c = CfgP()
        c.nx, c.ny, c.patch, c.channels, c.dim = j["nx"], j["ny"], j["patch"], j["channels"], j["dim"]
        c.layers, c.heads, c.meta_dim = j["layers"], j["heads"], j["meta_dim"]
        c.head_dim = self.w["head.to_cells.weight"].shape[0]
        c.ff_inner = self.w["blocks.0.ff1.weight"].shape[1]
        m = j.get("mask", "Full")
        if m == "Full":
            c.mask_kind, c.radius, c.dilation = 0, 0, 1
        else:
            c.mask_kind, c.radius, c.dilation = 1, m["Stencil"]["radius"], m["Stencil"]["dilation"]
        return c

    def phys(self) -> Phys:
        n = self.cfg["norm"]
        return phys(n["mean"], n["std"], n["depth"], self.cfg["g"], self.cfg.get("energy_r_max", 1.0))

    def forward(self, state: np.ndarray, meta: np.ndarray, f64: bool = False) -> np.ndarray:
        c = self.c_cfg()
        dt = np.float64 if f64 else np.float32
        ws = [np.ascontiguousarray(self.w[n], dtype=dt) for n in p_names(c.layers)]
        s = np.ascontiguousarray(state, dtype=dt).reshape(-1)
        m = np.ascontiguousarray(meta, dtype=dt).reshape(-1)
        out = np.zeros_like(s)
        fn = lib().orc_forward_p_f64 if f64 else lib().orc_forward_p_f32
        fn(ctypes.byref(c), _wptrs(ws), _ptr(s), _ptr(m), _ptr(out))
        return out

    def step(self, state, meta, f64: bool = False) -> np.ndarray:
        """cpu.rs:84-88: state + forward(state)."""
        dt = np.float64 if f64 else np.float32
        s = np.ascontiguousarray(state, dtype=dt).reshape(-1)
        out = self.forward(s, meta, f64)
        out += s
        return out


def mass(x: np.ndarray, n: int) -> float:
    x = np.ascontiguousarray(x)
    if x.dtype == np.float64:
        return lib().orc_mass_f64(_ptr(x), n)
    return lib().orc_mass_f32(_ptr(x), n)


def energy(x: np.ndarray, p: Phys, n: int) -> float:
    x = np.ascontiguousarray(x)
    if x.dtype == np.float64:
        return lib().orc_energy_f64(_ptr(x), ctypes.byref(p), n)
    return lib().orc_energy_f32(_ptr(x), ctypes.byref(p), n)


def project_mass(x: np.ndarray, target: float, n: int) -> None:
    assert x.flags.c_contiguous
    if x.dtype == np.float64:
        lib().orc_project_mass_f64(_ptr(x), target, n)
    else:
        lib().orc_project_mass_f32(_ptr(x), target, n)


def project_energy(x: np.ndarray, prev_energy: float, p: Phys, n: int) -> None:
    assert x.flags.c_contiguous
    if x.dtype == np.float64:
        lib().orc_project_energy_f64(_ptr(x), prev_energy, ctypes.byref(p), n)
    else:
        lib().orc_project_energy_f32(_ptr(x), prev_energy, ctypes.byref(p), n)


def mask_allowed(kind: int, radius: int, dilation: int, tokens_x: int, q: int, k: int) -> bool:
    return bool(lib().orc_mask_allowed(kind, radius, dilation, tokens_x, q, k))


def gelu(x: float) -> float:
    return lib().orc_gelu_f32(x)


# ---------------------------------------------------------------- form C

# bf16 rounding points of the trunk (orc_set_act_bf16)
ACT_BF16_LN1, ACT_BF16_QKV, ACT_BF16_ATTN, ACT_BF16_LN2, ACT_BF16_GELU = 1, 2, 4, 8, 16
ACT_BF16_ALL = 31


@dataclass
class FormC:
    """Form C model (per-cell space-time tokens): config dict + tensors.
    act_bf16 is a mask of trunk values rounded to bf16 (ACT_BF16_* below;
    ACT_BF16_ALL is the served bf16 form's storage, see orc_set_act_bf16)."""
    cfg: dict
    w: dict
    act_bf16: int = 0

    @staticmethod
    def load(dir_or_bytes, config_json: str | None = None) -> "FormC":
        if config_json is None:
            d = Path(dir_or_bytes)
            return FormC(json.loads((d / "config.json").read_text()), read_safetensors((d / "model.safetensors").read_bytes()))
        return FormC(json.loads(config_json), read_safetensors(dir_or_bytes))

    def c_cfg(self, T: int, ny: int, nx: int) -> CfgC:
        j = self.cfg
        c = CfgC()
        c.nx, c.ny, c.T, c.channels = nx, ny, T, 3
        c.dim, c.layers, c.heads, c.meta_dim = j["dim"], j["layers"], j["heads"], j["meta_dim"]
        c.ff_inner = self.w["blocks.0.ff1.weight"].shape[1]
        st = j["stencil"]
        c.rs, c.rt = st["radius"], st["time_radius"]
        return c

    def phys(self) -> Phys:
        n = self.cfg["norm"]
        return phys(n["mean"], n["std"], n["depth"], self.cfg["g"], self.cfg.get("energy_r_max", 1.0))

    def weights(self, f64: bool):
        dt = np.float64 if f64 else np.float32
        return [np.ascontiguousarray(self.w[n], dtype=dt) for n in c_names(self.cfg["layers"])]

    def forward(self, grid: np.ndarray, known: np.ndarray, meta: np.ndarray, f64: bool = False) -> np.ndarray:
        """grid [3,T,ny,nx] -> y [T*ny*nx, 3]."""
        _, T, ny, nx = grid.shape
        c = self.c_cfg(T, ny, nx)
        dt = np.float64 if f64 else np.float32
        ws = self.weights(f64)
        g = np.ascontiguousarray(grid, dtype=dt)
        k = np.ascontiguousarray(known, dtype=np.uint8)
        m = np.ascontiguousarray(meta, dtype=dt)
        y = np.zeros((T * ny * nx, 3), dtype=dt)
        fn = lib().orc_forward_c_f64 if f64 else lib().orc_forward_c_f32
        lib().orc_set_act_bf16(int(self.act_bf16))
        try:
            fn(ctypes.byref(c), _wptrs(ws), _ptr(g), _ptr(k), _ptr(m), _ptr(y))
        finally:
            lib().orc_set_act_bf16(0)
        return y

    def iterate(self, x0: np.ndarray, known: np.ndarray, meta: np.ndarray, max_iter: int, tol: float,
                f64: bool = False):
        """Returns (final grid, iterations, residual trace)."""
        _, T, ny, nx = x0.shape
        c = self.c_cfg(T, ny, nx)
        dt = np.float64 if f64 else np.float32
        ws = self.weights(f64)
        x = np.ascontiguousarray(x0, dtype=dt).copy()
        k = np.ascontiguousarray(known, dtype=np.uint8)
        m = np.ascontiguousarray(meta, dtype=dt)
        tr = np.zeros(max(max_iter, 1), dtype=dt)
        p = self.phys()
        fn = lib().orc_iterate_c_f64 if f64 else lib().orc_iterate_c_f32
        lib().orc_set_act_bf16(int(self.act_bf16))
        try:
            its = fn(ctypes.byref(c), _wptrs(ws), ctypes.byref(p), _ptr(x), _ptr(k), _ptr(m), max_iter, tol, _ptr(tr))
        finally:
            lib().orc_set_act_bf16(0)
        return x, its, tr[:its]

    def iterate_states(self, x0: np.ndarray, known: np.ndarray, meta: np.ndarray, max_iter: int, tol: float = 0.0,
                       f64: bool = False):
        """Like iterate, but returns every iterate: (states [its, 3, T, ny, nx],
        iterations, residual trace)."""
        _, T, ny, nx = x0.shape
        c = self.c_cfg(T, ny, nx)
        dt = np.float64 if f64 else np.float32
        ws = self.weights(f64)
        x = np.ascontiguousarray(x0, dtype=dt).copy()
        k = np.ascontiguousarray(known, dtype=np.uint8)
        m = np.ascontiguousarray(meta, dtype=dt)
        n = max(max_iter, 1)
        tr = np.zeros(n, dtype=dt)
        states = np.zeros((n,) + x0.shape, dtype=dt)
        p = self.phys()
        fn = lib().orc_iterate_c_states_f64 if f64 else lib().orc_iterate_c_states_f32
        lib().orc_set_act_bf16(int(self.act_bf16))
        try:
            its = fn(ctypes.byref(c), _wptrs(ws), ctypes.byref(p), _ptr(x), _ptr(k), _ptr(m), max_iter, tol,
                     _ptr(tr), _ptr(states))
        finally:
            lib().orc_set_act_bf16(0)
        return states[:its], its, tr[:its]


def fill_blank(grid: np.ndarray, known: np.ndarray) -> np.ndarray:
    _, T, ny, nx = grid.shape
    g = np.ascontiguousarray(grid, dtype=np.float32)
    k = np.ascontiguousarray(known, dtype=np.uint8)
    out = np.zeros_like(g)
    lib().orc_fill_blank(_ptr(g), _ptr(k), T, ny, nx, _ptr(out))
    return out


def slice_targets(x0: np.ndarray, p: Phys):
    _, T, ny, nx = x0.shape
    x = np.ascontiguousarray(x0, dtype=np.float32)
    mt = np.zeros(T)
    et = np.zeros(T)
    lib().orc_slice_targets_f64(_ptr(x), ctypes.byref(p), T, ny * nx, _ptr(mt), _ptr(et))
    return mt, et


def project_c(x: np.ndarray, known: np.ndarray, p: Phys, mass_t, eref_t) -> None:
    _, T, ny, nx = x.shape
    assert x.flags.c_contiguous
    k = np.ascontiguousarray(known, dtype=np.uint8)
    mt = np.ascontiguousarray(mass_t, dtype=np.float64)
    et = np.ascontiguousarray(eref_t, dtype=np.float64)
    fn = lib().orc_project_c_f64 if x.dtype == np.float64 else lib().orc_project_c_f32
    fn(_ptr(x), _ptr(k), ctypes.byref(p), T, ny * nx, _ptr(mt), _ptr(et))


def energy_violation(x: np.ndarray, p: Phys, eref_t) -> float:
    _, T, ny, nx = x.shape
    xx = np.ascontiguousarray(x, dtype=np.float32)
    et = np.ascontiguousarray(eref_t, dtype=np.float64)
    return lib().orc_energy_violation_c(_ptr(xx), ctypes.byref(p), T, ny * nx, _ptr(et))


def threads_default() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


# ---- shallow-water solver (sw_oracle.c) ----------------------------------------

class SwParams(ctypes.Structure):
    """crates/solvers/src/shallow_water.rs:16-29."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("dx", ctypes.c_float), ("dt", ctypes.c_float),
                ("g", ctypes.c_float), ("depth", ctypes.c_float), ("nu", ctypes.c_float)]


def sw_params(nx=64, ny=64, dx=1.0, dt=None, g=9.81, depth=1.0, nu=0.0) -> SwParams:
    """Params { .., ..Default::default() }: dt None takes the default's,
    with_cfl_dt(0.4) of dx 1, g 9.81, depth 1 (shallow_water.rs:31-58),
    evaluated in f32 as the reference does."""
    if dt is None:
        f = np.float32
        c = np.sqrt(f(9.81) * f(1.0), dtype=f)
        dt = f(0.4) * f(1.0) / (c * f(np.sqrt(2.0)))
    return SwParams(nx, ny, dx, float(dt), g, depth, nu)


def sw_steps(p: SwParams, h, u, v, steps: int):
    """ShallowWater::steps on copies; returns (h, u, v)."""
    h, u, v = (np.array(a, dtype=np.float32, order="C").copy() for a in (h, u, v))
    lib().orc_sw_steps(ctypes.byref(p), _ptr(h), _ptr(u), _ptr(v), int(steps))
    return h, u, v


def sw_add_drop(p: SwParams, h: np.ndarray, cx, cy, amp, sigma) -> None:
    """State::add_drop, in place on an f32 array."""
    assert h.dtype == np.float32 and h.flags.c_contiguous
    f = ctypes.c_float
    lib().orc_sw_add_drop(ctypes.byref(p), _ptr(h), f(cx), f(cy), f(amp), f(sigma))


def sw_trajectory(p: SwParams, h, u, v, ticks: int, drops=(), fields=None):
    """One dump-dataset trajectory: returns (traj [3, ticks+1, ny, nx], final h, u, v).
    drops: (tick, field) pairs into fields [k, ny, nx]."""
    h, u, v = (np.array(a, dtype=np.float32, order="C").copy() for a in (h, u, v))
    traj = np.empty((3, ticks + 1, p.ny, p.nx), dtype=np.float32)
    d = np.ascontiguousarray(np.array(drops, dtype=np.int32).reshape(-1, 2))
    fl = np.zeros((1, p.ny, p.nx), np.float32) if fields is None else np.ascontiguousarray(fields, dtype=np.float32)
    lib().orc_sw_trajectory(ctypes.byref(p), _ptr(h), _ptr(u), _ptr(v), int(ticks), _ptr(traj), len(d), _ptr(d),
                            _ptr(fl))
    return traj, h, u, v


def correlation(a, b) -> float:
    """metrics.rs:9-23."""
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    b = np.ascontiguousarray(b, dtype=np.float32).reshape(-1)
    f = lib().orc_correlation
    f.restype = ctypes.c_float
    return float(f(_ptr(a), _ptr(b), a.size))


# ---- field scores (metrics_oracle.c) -------------------------------------------

def _f32v(a):
    return np.ascontiguousarray(a, dtype=np.float32).reshape(-1)


def _fn(name, restype=ctypes.c_float):
    f = getattr(lib(), name)
    f.restype = restype
    return f


def peak(f) -> float:
    f = _f32v(f)
    return float(_fn("orc_peak")(_ptr(f), f.size))


def mean(f) -> float:
    f = _f32v(f)
    return float(_fn("orc_mean")(_ptr(f), f.size))


def std_dev(f) -> float:
    f = _f32v(f)
    return float(_fn("orc_std_dev")(_ptr(f), f.size))


def bias(model, solver) -> float:
    a, b = _f32v(model), _f32v(solver)
    return float(_fn("orc_bias")(_ptr(a), _ptr(b), a.size))


def radial_profile(f, nx: int, ny: int, cx: float, cy: float) -> np.ndarray:
    f = _f32v(f)
    out = np.zeros(nx + ny + 2, dtype=np.float64)
    fn = _fn("orc_radial_profile", ctypes.c_int)
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_void_p,
                   ctypes.c_int]
    k = fn(_ptr(f), nx, ny, cx, cy, _ptr(out), out.size)
    assert k > 0
    return out[:k].copy()


def xcorr_lag(model, solver, max_lag: int) -> float:
    a = np.ascontiguousarray(model, dtype=np.float64)
    b = np.ascontiguousarray(solver, dtype=np.float64)
    assert a.size == b.size
    return float(_fn("orc_xcorr_lag")(_ptr(a), _ptr(b), a.size, int(max_lag)))


def encode(h, u, v, mean3, std3, depth: float) -> np.ndarray:
    h, u, v = _f32v(h), _f32v(u), _f32v(v)
    m, s = _f32v(mean3), _f32v(std3)
    out = np.empty(3 * h.size, dtype=np.float32)
    fn = _fn("orc_encode", None)
    fn.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                           ctypes.c_void_p]
    fn(_ptr(h), _ptr(u), _ptr(v), h.size, _ptr(m), _ptr(s), depth, _ptr(out))
    return out


def decode_h(x, n: int, mean3, std3) -> np.ndarray:
    x = _f32v(x)
    m, s = _f32v(mean3), _f32v(std3)
    out = np.empty(n, dtype=np.float32)
    fn = _fn("orc_decode_h", None)
    fn(_ptr(x), n, _ptr(m), _ptr(s), _ptr(out))
    return out


def eval_scores(x, h, u, v, nx, ny, mean3, std3, depth, cx, cy, max_lag=8) -> dict:
    """eval.rs:158-172 for one trajectory at one horizon."""
    n = nx * ny
    x = _f32v(x)
    target = encode(h, u, v, mean3, std3, depth)
    ph, th = decode_h(x, n, mean3, std3), decode_h(target, n, mean3, std3)
    mp, tp = radial_profile(ph, nx, ny, cx, cy), radial_profile(th, nx, ny, cx, cy)
    return {"corr": correlation(x, target), "corr_h": correlation(x[:n], target[:n]),
            "amplitude_ratio": float(np.float32(peak(ph)) / np.float32(peak(th))),
            "lag": xcorr_lag(mp, tp, max_lag), "bias": bias(ph, th)}


# ---- excess-energy band (excess_oracle.c) --------------------------------------

def _f64v(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


def excess_terms(x, nx: int, ny: int, p: Phys, dt_dx: float):
    """project.py:87-101 for one grid x [3, ny, nx] (f64): (a, q)."""
    xx = _f64v(x)
    a, q = ctypes.c_double(), ctypes.c_double()
    fn = lib().orc_excess_terms
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                   ctypes.c_void_p]
    fn(_ptr(xx), nx, ny, ctypes.byref(p), dt_dx, ctypes.byref(a), ctypes.byref(q))
    return a.value, q.value


def excess_energy(x, nx: int, ny: int, p: Phys, dt_dx: float) -> float:
    a, q = excess_terms(x, nx, ny, p, dt_dx)
    return a + q


def project_excess_energy(x, nx: int, ny: int, p: Phys, dt_dx: float, prev: float, r_min: float, r_max: float):
    """project.py:104-140 on a copy (f64); returns (x', s)."""
    xx = _f64v(x).copy()
    fn = lib().orc_project_excess_energy
    fn.restype = ctypes.c_double
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                   ctypes.c_double, ctypes.c_double]
    s = fn(_ptr(xx), nx, ny, ctypes.byref(p), dt_dx, prev, r_min, r_max)
    return xx, s
