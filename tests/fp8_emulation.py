"""Numpy restatement of the FP8 form's arithmetic, for its GPU parity test
(and, with bf16_act, of the bf16 form's storage rounding):
the cell_st forward of the dense f64 restatement in test_oracle.py with the
four block products taken on E4M3 codes. Weights: per-tensor scale
amax / 448 (f32), round-to-nearest-even codes. Inputs (the LN1 output, the
attention output, the LN2 output, the GELU output): each row's (token's)
scale max|A[m, :]| / 448 (f32; 1 for a row of zeros), as libj2k measures it
on the device before each product, values clamped to +-448 before
rounding. Products are exact; sums in f64."""

import math

import numpy as np

from jacobi2000_amd import checkpoint as ck


def quantize(x, scale):
    x = np.clip(np.asarray(x, np.float32) / np.float32(scale), -448, 448)
    return ck.e4m3_decode(ck.e4m3_encode(x)).astype(np.float64)


def weight_codes(W):
    W = np.asarray(W, np.float32)
    sc = np.float32(np.abs(W).max()) / np.float32(448)
    return ck.e4m3_decode(ck.e4m3_encode((W / sc).astype(np.float32))).astype(np.float64), float(sc)


def bf16(a):
    """Round to bf16 through f32, nearest even."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def cell_forward(weights, cfg, grid, known, meta, fp8=True, bf16_act=False):
    """y = f(x) [T*ny*nx, 3] of the cell_st form; weights name -> array.
    bf16_act (with fp8 False) rounds what the served bf16 form stores as
    bf16: the LN outputs, q, k, v, the attention output, the GELU output."""
    w = {k: np.asarray(v, np.float64) for k, v in weights.items()}
    _, T, ny, nx = grid.shape
    t, y, x = np.meshgrid(np.arange(T), np.arange(ny), np.arange(nx), indexing="ij")
    t, y, x = t.reshape(-1), y.reshape(-1), x.reshape(-1)
    st = cfg["stencil"]
    dy = np.abs(y[:, None] - y[None, :])
    dx = np.abs(x[:, None] - x[None, :])
    dy, dx = np.minimum(dy, ny - dy), np.minimum(dx, nx - dx)
    allowed = (dy <= st["radius"]) & (dx <= st["radius"]) & (np.abs(t[:, None] - t[None, :]) <= st["time_radius"])
    g = np.asarray(grid, np.float64)
    feat = np.stack([g[0].reshape(-1), g[1].reshape(-1), g[2].reshape(-1), known.reshape(-1).astype(np.float64)], 1)
    X = feat @ w["embed.weight"] + w["embed.bias"] + (np.asarray(meta, np.float64) @ w["meta.weight"] + w["meta.bias"])

    def ln(z, gm, b):
        mu = z.mean(1, keepdims=True)
        var = ((z - mu) ** 2).mean(1, keepdims=True)
        return (z - mu) / np.sqrt(var + 1e-5) * gm + b

    H, d = cfg["heads"], cfg["dim"]
    hd = d // H
    gelu = np.vectorize(lambda v: 0.5 * v * (1 + math.erf(v / math.sqrt(2))))
    f32 = np.float32
    for i in range(cfg["layers"]):
        p = lambda s: w[f"blocks.{i}.{s}"]

        def lin(a, name):
            W, b = p(name + ".weight"), p(name + ".bias")
            if not fp8:
                return a @ W + b
            Wq, sw = weight_codes(W)
            amax = np.abs(np.asarray(a, np.float32)).max(1, keepdims=True)
            sa = np.where(amax > 0, amax / f32(448), f32(1)).astype(np.float32)
            return (quantize(a, sa) @ Wq) * (sa.astype(np.float64) * sw) + b

        r = bf16 if bf16_act else (lambda a: a)
        h = r(ln(X, p("norm1.gamma"), p("norm1.beta")))
        q, k, v = (r(lin(h, n)) for n in "qkv")
        out = np.zeros_like(X)
        for hh in range(H):
            sl = slice(hh * hd, (hh + 1) * hd)
            s = q[:, sl] @ k[:, sl].T / math.sqrt(hd)
            s = np.where(allowed, s, -np.inf)
            s = np.exp(s - s.max(1, keepdims=True))
            out[:, sl] = (s / s.sum(1, keepdims=True)) @ v[:, sl]
        X = X + lin(r(out), "proj")
        h = r(ln(X, p("norm2.gamma"), p("norm2.beta")))
        X = X + lin(r(gelu(lin(h, "ff1"))), "ff2")
    return ln(X, w["norm.gamma"], w["norm.beta"]) @ w["readout.weight"] + w["readout.bias"]
