"""Checkpoints and synthetic inputs in the engine's exchange format.

The format is the reference's: `model.safetensors` (F32 tensors, names as in
docs/pytorch-training.md:25-44) plus `config.json` (crates/engine/src/config.rs:13-41).
The per-cell space-time form ("cell_st") adds optional keys that the
reference ignores (config.rs reads a fixed field set and skips unknown keys):

    "form": "cell_st"
    "stencil": {"radius": rs, "time_radius": rt}   space periodic, time clipped
    "max_iter": 32, "tol": 1e-4                     fixed-point controls

and its tensors are embed.weight [4, dim] (features h, u, v, known),
embed.bias, meta.*, blocks.<i>.* (the reference block names), norm.*,
readout.weight [dim, 3], readout.bias [3]. There is no `pos` tensor, so one
checkpoint runs on any grid size.

Random initialisation: numpy's PCG64 generator (`numpy.random.default_rng`)
seeded with `seed`, drawing N(0, std) for every linear weight and bias in the
sorted order of tensor names; LayerNorm gamma = 1, beta = 0.
"""

from __future__ import annotations

import json
import struct

import numpy as np

BLOCK_LINEARS = ("q", "k", "v", "proj", "ff1", "ff2")


def write_safetensors(tensors: dict[str, np.ndarray], dtypes: dict[str, str] | None = None) -> bytes:
    """Byte layout of crates/engine/src/weights.rs:29-48 (8-byte LE header
    length, JSON header, LE buffer), tensors in sorted name order. F32 unless
    `dtypes` names "BF16" (the array is rounded to nearest even) or
    "F8_E4M3" (the array holds the uint8 codes) for a tensor."""
    header, off, blobs = {}, 0, []
    for name in sorted(tensors):
        dt = (dtypes or {}).get(name, "F32")
        if dt == "F32":
            a = np.ascontiguousarray(tensors[name], dtype="<f4")
        elif dt == "BF16":
            a = f32_to_bf16_bits(np.asarray(tensors[name], dtype=np.float32)).astype("<u2")
        elif dt == "F8_E4M3":
            a = np.ascontiguousarray(tensors[name], dtype=np.uint8)
        else:
            raise ValueError(f"dtype {dt}")
        header[name] = {"dtype": dt, "shape": list(a.shape), "data_offsets": [off, off + a.nbytes]}
        off += a.nbytes
        blobs.append(a.tobytes())
    h = json.dumps(header, separators=(",", ":")).encode()
    h += b" " * ((8 - len(h) % 8) % 8)
    return struct.pack("<Q", len(h)) + h + b"".join(blobs)


def f32_to_bf16_bits(x: np.ndarray) -> np.ndarray:
    """bf16 bit patterns of f32 values, rounding to nearest even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def e4m3_decode(codes: np.ndarray) -> np.ndarray:
    """OCP FP8 E4M3FN codes -> float32: 1 sign, 4 exponent (bias 7), 3
    mantissa bits; exponent 0 is subnormal (m/8 * 2^-6); S.1111.111 is NaN."""
    c = np.asarray(codes, dtype=np.uint8).astype(np.int32)
    e, m = (c >> 3) & 15, c & 7
    mag = np.where(e == 0, np.ldexp(m.astype(np.float64), -9), np.ldexp(1.0 + m / 8.0, e - 7))
    mag = np.where((e == 15) & (m == 7), np.nan, mag)
    return np.where(c >> 7, -mag, mag).astype(np.float32)


_E4M3_POS = e4m3_decode(np.arange(0x7F, dtype=np.uint8)).astype(np.float64)  # codes 0..126, ascending


def e4m3_encode(x: np.ndarray) -> np.ndarray:
    """float -> OCP E4M3FN codes, round to nearest (ties to the even code),
    saturating to +-448; NaN -> 0x7F."""
    x = np.asarray(x, dtype=np.float64)
    a = np.minimum(np.abs(x), 448.0)
    lo = np.clip(np.searchsorted(_E4M3_POS, a, side="right") - 1, 0, 0x7E)
    hi = np.minimum(lo + 1, 0x7E)
    dlo, dhi = a - _E4M3_POS[lo], _E4M3_POS[hi] - a
    code = np.where((dhi < dlo) | ((dhi == dlo) & (lo % 2 == 1)), hi, lo).astype(np.uint8)
    code = np.where(np.signbit(x), code | 0x80, code).astype(np.uint8)
    return np.where(np.isnan(x), np.uint8(0x7F), code).astype(np.uint8)


def convert_checkpoint(st: bytes, dtype: str) -> bytes:
    """A copy of a checkpoint with the block weight matrices (q, k, v, proj,
    ff1, ff2) stored as BF16, or as F8_E4M3 codes with a per-tensor scale
    `<name>.scale` (amax / 448). The extension of the reference format that
    libj2k reads; every other tensor stays F32."""
    if dtype not in ("BF16", "F8_E4M3"):
        raise ValueError("dtype must be BF16 or F8_E4M3")
    t = read_safetensors(st)
    out, dtypes = {}, {}
    for name, a in t.items():
        parts = name.split(".")
        if len(parts) == 4 and parts[0] == "blocks" and parts[2] in BLOCK_LINEARS and parts[3] == "weight":
            if dtype == "BF16":
                out[name], dtypes[name] = a, "BF16"
            else:
                # f32 arithmetic, as libj2k quantizes an F32 checkpoint at load
                amax = np.float32(np.abs(a).max())
                scale = amax / np.float32(448.0) if amax > 0 else np.float32(1.0)
                out[name], dtypes[name] = e4m3_encode((a.astype(np.float32) / scale).astype(np.float32)), "F8_E4M3"
                out[name + ".scale"] = np.array([scale], np.float32)
        else:
            out[name] = a
    return write_safetensors(out, dtypes)


def read_safetensors(data: bytes) -> dict[str, np.ndarray]:
    """Tensors of a checkpoint as float32 (BF16 widened, F8_E4M3 decoded
    times `<name>.scale`; the scale tensors are dropped)."""
    n = struct.unpack("<Q", data[:8])[0]
    header = json.loads(data[8:8 + n])
    raw, out = {}, {}
    for name, v in header.items():
        if name == "__metadata__":
            continue
        a, b = v["data_offsets"]
        buf = data[8 + n + a:8 + n + b]
        if v["dtype"] == "F32":
            raw[name] = np.frombuffer(buf, dtype="<f4").reshape(v["shape"]).copy()
        elif v["dtype"] == "BF16":
            bits = np.frombuffer(buf, dtype="<u2").astype(np.uint32) << 16
            raw[name] = bits.view(np.float32).reshape(v["shape"]).copy()
        elif v["dtype"] == "F8_E4M3":
            raw[name] = ("f8", np.frombuffer(buf, dtype=np.uint8).reshape(v["shape"]).copy())
        else:
            raise ValueError(f"{name}: dtype {v['dtype']}")
    for name, a in raw.items():
        if isinstance(a, tuple):
            out[name] = e4m3_decode(a[1]) * raw[name + ".scale"].reshape(-1)[0]
        elif not (name.endswith(".scale") and isinstance(raw.get(name[:-6]), tuple)):
            out[name] = a
    return out


def cell_st_shapes(dim: int, layers: int, ff: int, in_features: int = 4, meta_dim: int = 5,
                   channels: int = 3) -> dict[str, tuple]:
    s = {"embed.weight": (in_features, dim), "embed.bias": (dim,), "meta.weight": (meta_dim, dim), "meta.bias": (dim,),
         "norm.gamma": (dim,), "norm.beta": (dim,), "readout.weight": (dim, channels), "readout.bias": (channels,)}
    for i in range(layers):
        p = f"blocks.{i}."
        s[p + "norm1.gamma"] = (dim,)
        s[p + "norm1.beta"] = (dim,)
        s[p + "norm2.gamma"] = (dim,)
        s[p + "norm2.beta"] = (dim,)
        for n, (a, b) in {"q": (dim, dim), "k": (dim, dim), "v": (dim, dim), "proj": (dim, dim),
                          "ff1": (dim, ff), "ff2": (ff, dim)}.items():
            s[p + n + ".weight"] = (a, b)
            s[p + n + ".bias"] = (b,)
    return s


def random_cell_st(dim: int = 128, layers: int = 4, heads: int = 4, ff: int = 512, seed: int = 0,
                   std: float = 0.02, radius: int = 1, time_radius: int = 1, max_iter: int = 32,
                   tol: float = 1e-4, r_max: float = 1.0) -> tuple[bytes, str]:
    """A random-init cell_st checkpoint: (model.safetensors bytes, config.json text).
    Identity normalizer with depth 1 (normalized h = h - 1), g = 9.81."""
    rng = np.random.default_rng(seed)
    shapes = cell_st_shapes(dim, layers, ff)
    t = {}
    for name in sorted(shapes):
        shp = shapes[name]
        if name.endswith(".gamma"):
            t[name] = np.ones(shp, np.float32)
        elif name.endswith(".beta"):
            t[name] = np.zeros(shp, np.float32)
        else:
            t[name] = rng.normal(0.0, std, size=shp).astype(np.float32)
    cfg = {"form": "cell_st", "dim": dim, "layers": layers, "heads": heads, "channels": 3, "meta_dim": 5,
           "ff_mult": ff // dim, "norm": {"mean": [0.0, 0.0, 0.0], "std": [1.0, 1.0, 1.0], "depth": 1.0},
           "dx": 1.0, "dt": 0.09, "g": 9.81, "depth": 1.0, "nu": 1e-3, "energy_r_max": r_max,
           "stencil": {"radius": radius, "time_radius": time_radius}, "max_iter": max_iter, "tol": tol}
    return write_safetensors(t), json.dumps(cfg, indent=2)


def meta_of(dx: float, dt: float, g: float, depth: float, nu: float) -> np.ndarray:
    """crates/engine/src/solver_meta.rs:9-11, in f32 as the reference computes it."""
    f = np.float32
    return np.array([np.log(f(dx)), np.log(f(dt)), np.log(f(g)), np.log(f(depth)), np.log(f(nu) + f(1e-6))],
                    dtype=np.float32)


def add_drop(h: np.ndarray, cx: float, cy: float, amp: float, sigma: float) -> None:
    """crates/solvers/src/shallow_water.rs:102-111 on an [ny, nx] field, f32."""
    ny, nx = h.shape
    f = np.float32
    inv = f(1.0) / (f(2.0) * f(sigma) * f(sigma))
    ys, xs = np.meshgrid(np.arange(ny, dtype=np.float32), np.arange(nx, dtype=np.float32), indexing="ij")
    dx = xs - f(cx)
    dy = ys - f(cy)
    h += (f(amp) * np.exp(-(dx * dx + dy * dy) * inv)).astype(np.float32)


def drops_grid(T: int, ny: int, nx: int, drops: int, seed: int, centred: bool = False,
               amp_range=(0.05, 0.15), sigma_range=(2.0, 5.0), amp: float = 0.1, sigma: float = 3.0):
    """Synthetic request: normalized grid [3, T, ny, nx] with every slice equal
    to a flat state (h = depth) plus Gaussian drops, u = v = 0. With `centred`,
    one drop of (amp, sigma) at the grid centre. Identity normalizer, depth 1,
    so the normalized h is the surface anomaly."""
    h = np.zeros((ny, nx), np.float32)
    if centred:
        add_drop(h, nx / 2.0, ny / 2.0, amp, sigma)
    else:
        rng = np.random.default_rng(seed)
        for _ in range(drops):
            cx, cy = rng.uniform(0, nx), rng.uniform(0, ny)
            a, s = rng.uniform(*amp_range), rng.uniform(*sigma_range)
            add_drop(h, cx, cy, a, s)
    g = np.zeros((3, T, ny, nx), np.float32)
    g[0] = h[None]
    return g


def moving_drops_grid(T: int, ny: int, nx: int, drops: int, seed: int, amp_range=(0.05, 0.15),
                      sigma_range=(2.0, 5.0), speed: float = 0.75, flow: float = 0.02) -> np.ndarray:
    """Synthetic request whose slices differ: Gaussian drops (amplitude and
    width drawn as in drops_grid) whose centres move `speed` cells per slice
    in seeded directions, wrapping periodically, plus a weak smooth flow
    u = flow sin(2 pi (y/ny + t/T)), v = flow cos(2 pi (x/nx + t/T)).
    Normalized grid [3, T, ny, nx] (identity normalizer, depth 1)."""
    rng = np.random.default_rng(seed)
    f = np.float32
    ys, xs = np.meshgrid(np.arange(ny, dtype=np.float64), np.arange(nx, dtype=np.float64), indexing="ij")
    g = np.zeros((3, T, ny, nx), np.float32)
    for _ in range(drops):
        cx, cy = rng.uniform(0, nx), rng.uniform(0, ny)
        a, s = rng.uniform(*amp_range), rng.uniform(*sigma_range)
        th = rng.uniform(0, 2 * np.pi)
        for t in range(T):
            px, py = cx + speed * t * np.cos(th), cy + speed * t * np.sin(th)
            dx = (xs - px + nx / 2.0) % nx - nx / 2.0  # nearest periodic image
            dy = (ys - py + ny / 2.0) % ny - ny / 2.0
            g[0, t] += (a * np.exp(-(dx * dx + dy * dy) / (2.0 * s * s))).astype(f)
    for t in range(T):
        g[1, t] = (flow * np.sin(2 * np.pi * (ys / ny + t / T))).astype(f)
        g[2, t] = (flow * np.cos(2 * np.pi * (xs / nx + t / T))).astype(f)
    return g


def mask_of(kind: str, T: int, ny: int, nx: int, seed: int = 0, frac: float = 0.02) -> np.ndarray:
    """Known-cell mask [T, ny, nx] (uint8): forward = slice 0 known, rewind =
    last slice known, reconstruction = `frac` of the (y, x) cells drawn
    uniformly at random, each known in every slice."""
    k = np.zeros((T, ny, nx), np.uint8)
    if kind == "forward":
        k[0] = 1
    elif kind == "rewind":
        k[T - 1] = 1
    elif kind == "reconstruction":
        rng = np.random.default_rng(seed)
        n = ny * nx
        idx = rng.choice(n, size=max(1, int(round(frac * n))), replace=False)
        k.reshape(T, n)[:, idx] = 1
    else:
        raise ValueError(f"unknown mask kind `{kind}` (want forward | rewind | reconstruction)")
    return k


def patch2d_shapes(nx: int, ny: int, patch: int, dim: int, layers: int, ff: int, head_dim: int, channels: int = 3,
                   meta_dim: int = 5) -> dict[str, tuple]:
    """Tensor shapes of the reference exchange format (docs/pytorch-training.md:25-44)."""
    tokens = (nx // patch) * (ny // patch)
    s = {"embed.weight": (channels * patch * patch, dim), "embed.bias": (dim,), "meta.weight": (meta_dim, dim),
         "meta.bias": (dim,), "pos": (1, tokens, dim), "norm.gamma": (dim,), "norm.beta": (dim,),
         "head.token_mix.weight": (dim, dim, 3, 3), "head.token_mix.bias": (dim,),
         "head.to_cells.weight": (head_dim, dim, 1, 1), "head.to_cells.bias": (head_dim,),
         "head.conv1.weight": (head_dim, head_dim + channels, 3, 3), "head.conv1.bias": (head_dim,),
         "head.conv2.weight": (channels, head_dim, 3, 3), "head.conv2.bias": (channels,)}
    for k, v in cell_st_shapes(dim, layers, ff).items():
        if k.startswith("blocks."):
            s[k] = v
    return s


def random_patch2d(nx: int = 64, ny: int = 64, patch: int = 4, dim: int = 128, layers: int = 4, heads: int = 4,
                   ff: int = 512, head_dim: int = 32, seed: int = 0, std: float = 0.02,
                   mask=None, r_max: float = 1.0001) -> tuple[bytes, str]:
    """A random-init checkpoint of the reference patch transformer (model20
    shape by default), same PCG64 N(0, std) rule as random_cell_st."""
    rng = np.random.default_rng(seed)
    shapes = patch2d_shapes(nx, ny, patch, dim, layers, ff, head_dim)
    t = {}
    for name in sorted(shapes):
        shp = shapes[name]
        if name.endswith(".gamma"):
            t[name] = np.ones(shp, np.float32)
        elif name.endswith(".beta"):
            t[name] = np.zeros(shp, np.float32)
        else:
            t[name] = rng.normal(0.0, std, size=shp).astype(np.float32)
    cfg = {"nx": nx, "ny": ny, "patch": patch, "channels": 3, "dim": dim, "layers": layers, "heads": heads,
           "ff_mult": ff // dim, "meta_dim": 5,
           "norm": {"mean": [0.0, 0.0, 0.0], "std": [1.0, 1.0, 1.0], "depth": 1.0},
           "dx": 1.0, "dt": 0.09, "g": 9.81, "depth": 1.0, "nu": 1e-3, "energy_r_max": r_max,
           "mask": mask if mask is not None else {"Stencil": {"radius": 1, "dilation": 1}}}
    return write_safetensors(t), json.dumps(cfg, indent=2)
