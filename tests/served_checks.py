"""Shared inputs, error measures and tolerances of the served-path parity
tests (test_gpu_served.py) and of their planted-mistake checks on the CPU
(test_oracle_teeth.py).

Two checkpoints of the model20 shape (d = 128, 4 layers, 4 heads x 32,
ff 512, 27-key stencil):
  - model20: N(0, 0.02) from seed 0, energy_r_max 1.0, the configured
    workload. Its output is nearly a per-slice constant and the energy
    projection pulls the iterate back to the anchor, so only the raw model
    output y shows whether the trunk is right.
  - teeth: N(0, 0.1) from seed 5, energy_r_max 1.05. Its output depends on
    the input, and the energy bound leaves room, so every field of every
    iterate carries the model's work.

The served forms multiply bf16 weights. Their arithmetic is pinned against
the f64 oracle run on those same bf16-rounded weights (bf16x2 keeps f32
activations as bf16 pairs, about 16 mantissa bits, so it must sit close to
that oracle); their distance to the exact-weight oracle is the precision
cost, reported beside it.
"""

from __future__ import annotations

import json

import numpy as np

from jacobi2000_amd import checkpoint as ck
from oracle import oracle as O

META = ck.meta_of(1.0, 0.09, 9.81, 1.0, 1e-3)
KINDS = ("forward", "rewind", "reconstruction")
BLOCK_MATRICES = ("q", "k", "v", "proj", "ff1", "ff2")


def model20():
    return ck.random_cell_st(seed=0)


def teeth():
    return ck.random_cell_st(seed=5, std=0.1, r_max=1.05)


def request(kind: str, T: int, ny: int, nx: int, seed: int = 11, moving: bool = True):
    """(x0 with blank cells filled, known mask): five seeded drops; with
    `moving` the drops travel and a weak flow varies over the slices, so the
    slices differ and the time stencil matters."""
    g = ck.moving_drops_grid(T, ny, nx, 5, seed) if moving else ck.drops_grid(T, ny, nx, 5, seed)
    k = ck.mask_of(kind, T, ny, nx, seed=1)
    return O.fill_blank(g, k), k


def bf16_round(a: np.ndarray) -> np.ndarray:
    return (ck.f32_to_bf16_bits(a).astype(np.uint32) << 16).view(np.float32)


def bf16_weights(cpu: O.FormC) -> O.FormC:
    """The oracle model with the block matrices rounded to bf16 (nearest
    even), the weights the served forms multiply."""
    w = {}
    for name, a in cpu.w.items():
        p = name.split(".")
        quant = len(p) == 4 and p[0] == "blocks" and p[2] in BLOCK_MATRICES and p[3] == "weight"
        w[name] = bf16_round(a) if quant else a
    return O.FormC(cpu.cfg, w)


def bf16_form(cpu: O.FormC) -> O.FormC:
    """The oracle with the bf16 form's arithmetic: bf16 block matrices and
    the trunk's stored activations rounded to bf16."""
    return O.FormC(cpu.cfg, bf16_weights(cpu).w, act_bf16=O.ACT_BF16_ALL)


PLANTS = ("readout_zero", "last_attention_dropped", "first_mlp_dropped", "space_radius_0", "time_radius_0")


def planted(cpu: O.FormC, what: str) -> O.FormC:
    """The oracle with one plausible mistake: the readout zeroed (f = bias),
    the last block's attention output dropped, the first block's MLP output
    dropped, the space stencil collapsed to one cell, or the time stencil to
    one slice."""
    w = {k: v.copy() for k, v in cpu.w.items()}
    cfg = json.loads(json.dumps(cpu.cfg))
    last = cfg["layers"] - 1
    if what == "readout_zero":
        w["readout.weight"][:] = 0
    elif what == "last_attention_dropped":
        w[f"blocks.{last}.proj.weight"][:] = 0
        w[f"blocks.{last}.proj.bias"][:] = 0
    elif what == "first_mlp_dropped":
        w["blocks.0.ff2.weight"][:] = 0
        w["blocks.0.ff2.bias"][:] = 0
    elif what == "space_radius_0":
        cfg["stencil"]["radius"] = 0
    elif what == "time_radius_0":
        cfg["stencil"]["time_radius"] = 0
    else:
        raise ValueError(what)
    return O.FormC(cfg, w)


def rel_var(y: np.ndarray, ref: np.ndarray) -> np.ndarray:
    """Per field of a model output [M, 3]: ||y - ref|| over the norm of
    ref's variation about its own mean (the mean is nearly all of model20's
    output, so a plain relative norm would hide a wrong trunk)."""
    a, b = np.asarray(y, np.float64).T, np.asarray(ref, np.float64).T
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b - b.mean(1, keepdims=True), axis=1)


def rel_inc(x: np.ndarray, ref: np.ndarray, x0: np.ndarray) -> np.ndarray:
    """Per field of an iterate [3, T, ny, nx]: ||x - ref|| over ||ref - x0||,
    the error against the part the iterations made."""
    a = np.asarray(x, np.float64).reshape(3, -1)
    b = np.asarray(ref, np.float64).reshape(3, -1)
    c = np.asarray(x0, np.float64).reshape(3, -1)
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b - c, axis=1), 1e-300)


def rel_l2(x: np.ndarray, ref: np.ndarray) -> float:
    a, b = np.asarray(x, np.float64), np.asarray(ref, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def own_oracle(cpu: O.FormC, precision: str) -> O.FormC:
    """The oracle a form is compared with: exact weights for f32, bf16
    weights for bf16x2, bf16 weights and bf16 storage rounding for bf16."""
    return {"f32": cpu, "bf16x2": bf16_weights(cpu), "bf16": bf16_form(cpu)}[precision]


# Tolerances. Raw output y against the f64 oracle on the form's own
# arithmetic (rel_var per field). Pure bf16 storage is chaotic under
# rounding: the bf16-form oracle itself moves by 5-20% of the output's
# variation between its f32 and f64 builds (bf16 rounding flips), so that
# form is held to a bound of that size; its kernels are the bf16x2 form's
# with the operand split left out, and bf16x2 is held to 5e-3.
Y_TOL = {"f32": 1e-3, "bf16x2": 5e-3, "bf16": 0.5}
# Iterates against the oracle on the form's own arithmetic (rel_inc per
# field); the f32 form is also held to 1e-4 max abs per field per iteration
# against the f32 oracle. Pure bf16 iterates are held to 2e-2 relative L2 of
# the state (rel_inc of its h increment, the small part left after the mass
# shift, is below its rounding noise).
INC_TOL = {"f32": 1e-3, "bf16x2": 5e-3}
BF16_STATE_TOL = 2e-2
F32_MAX_ABS = 1e-4
