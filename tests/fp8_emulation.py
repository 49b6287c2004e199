"""Numpy restatement of the FP8 form's arithmetic, for its GPU parity test
(and, with bf16_act, of the bf16 form's storage rounding):
the cell_st forward of the dense f64 restatement in test_oracle.py with the
four block products taken on E4M3 codes. Weights: per-tensor scale
amax / 448 (f32), round-to-nearest-even codes. Inputs: the static
per-tensor scales libj2k derives from the weights (LN output bound
max|gamma| sqrt(d-1) + max|beta|; a linear output bound bound_in *
max_n sum_k |Wq[k, n]| + |b_n| over the quantized weights; GELU bounded by
max(bound, 0.17)), values
clamped to +-448 before rounding. Products are exact; sums in f64."""

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
        bound = lambda gm, b: float(f32(np.abs(gm).max()) * f32(math.sqrt(d - 1)) + f32(np.abs(b).max()))
        bh1, bh2 = bound(p("norm1.gamma"), p("norm1.beta")), bound(p("norm2.gamma"), p("norm2.beta"))
        wv, sv = weight_codes(p("v.weight"))
        w1, s1 = weight_codes(p("ff1.weight"))
        bo = float((bh1 * np.abs(wv * sv).sum(0) + np.abs(p("v.bias"))).max())
        bz = float((bh2 * np.abs(w1 * s1).sum(0) + np.abs(p("ff1.bias"))).max())

        def lin(a, name, sa):
            W, b = p(name + ".weight"), p(name + ".bias")
            if not fp8:
                return a @ W + b
            Wq, sw = weight_codes(W)
            return (quantize(a, sa) @ Wq) * sa * sw + b

        r = bf16 if bf16_act else (lambda a: a)
        h = r(ln(X, p("norm1.gamma"), p("norm1.beta")))
        q, k, v = (r(lin(h, n, bh1 / 448)) for n in "qkv")
        out = np.zeros_like(X)
        for hh in range(H):
            sl = slice(hh * hd, (hh + 1) * hd)
            s = q[:, sl] @ k[:, sl].T / math.sqrt(hd)
            s = np.where(allowed, s, -np.inf)
            s = np.exp(s - s.max(1, keepdims=True))
            out[:, sl] = (s / s.sum(1, keepdims=True)) @ v[:, sl]
        X = X + lin(r(out), "proj", bo / 448)
        h = r(ln(X, p("norm2.gamma"), p("norm2.beta")))
        X = X + lin(r(gelu(lin(h, "ff1", bh2 / 448))), "ff2", max(bz, 0.17) / 448)
    return ln(X, w["norm.gamma"], w["norm.beta"]) @ w["readout.weight"] + w["readout.bias"]
