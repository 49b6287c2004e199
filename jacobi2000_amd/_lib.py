"""ctypes binding of libj2k.so (include/j2k.h).

The library is the product: every call here goes to HIP code on the GPU. A
missing library is an error, never a fallback.
"""

from __future__ import annotations

import ctypes
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libj2k.so"

OK = 0
ERR_NO_ADAPTER, ERR_NO_DEVICE, ERR_LIMIT, ERR_MAP, ERR_POLL = 1, 2, 3, 4, 5
ERR_BAD_CONFIG, ERR_MISSING_TENSOR, ERR_BAD_DTYPE, ERR_SHAPE, ERR_INVALID, ERR_OOM = 6, 7, 8, 9, 10, 11
PENDING = 12
PREC_F32, PREC_BF16, PREC_BF16X2, PREC_FP8 = 0, 1, 2, 3
FORM_PATCH2D, FORM_CELL_ST = 0, 1

_NAMES = {1: "NoAdapter", 2: "NoDevice", 3: "Limit", 4: "Map", 5: "Poll", 6: "BadConfig", 7: "MissingTensor",
          8: "BadDType", 9: "ShapeMismatch", 10: "Invalid", 11: "OutOfMemory", 12: "Pending"}


class J2KError(RuntimeError):
    """A failed libj2k call: `.code` is the C-ABI status, `.kind` its name."""

    def __init__(self, code: int, msg: str):
        self.code = code
        self.kind = _NAMES.get(code, str(code))
        super().__init__(f"{self.kind}: {msg}")


class Config(ctypes.Structure):
    _fields_ = ([(f, ctypes.c_int) for f in ("form", "precision", "nx", "ny", "patch", "channels", "dim", "layers",
                                            "heads", "head_dim", "ff_inner", "meta_dim", "tokens_x", "tokens_y",
                                            "tokens", "patch_dim", "mask_kind", "radius", "dilation",
                                            "time_radius")]
                + [("norm_mean", ctypes.c_float * 3), ("norm_std", ctypes.c_float * 3), ("norm_depth", ctypes.c_float)]
                + [(f, ctypes.c_float) for f in ("dx", "dt", "g", "depth", "nu", "energy_r_max")]
                + [("max_iter", ctypes.c_int), ("tol", ctypes.c_float), ("stall_stop", ctypes.c_int)]
                + [("excess_r", ctypes.c_float), ("excess_r_min", ctypes.c_float)])


class Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("iters", ctypes.c_int), ("converged", ctypes.c_int),
                ("stalled", ctypes.c_int), ("resid", ctypes.c_float), ("violation", ctypes.c_double), ("submit_s", ctypes.c_double),
                ("start_s", ctypes.c_double), ("done_s", ctypes.c_double)]


class SwParams(ctypes.Structure):
    """j2k_sw_params: crates/solvers/src/shallow_water.rs:16-29."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int)] + [(f, ctypes.c_float) for f in ("dx", "dt", "g", "depth",
                                                                                               "nu")]


class Score(ctypes.Structure):
    """j2k_eval_score: eval-metrics' scores (crates/engine-cli/src/eval.rs:158-172)."""
    _fields_ = [(f, ctypes.c_float) for f in ("corr", "corr_h", "amplitude_ratio", "lag", "bias")]


class SwDrop(ctypes.Structure):
    _fields_ = [("tick", ctypes.c_int), ("field", ctypes.c_int)]


P = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_float
D = ctypes.c_double
PP = ctypes.POINTER(ctypes.c_void_p)

# name: argtypes (every function returns int status unless noted)
PROTOS = {
    "j2k_device_count": [ctypes.POINTER(I)],
    "j2k_model_create": [P, ctypes.c_size_t, ctypes.c_char_p, I, I, PP],
    "j2k_model_destroy": [P],
    "j2k_model_config": [P, ctypes.POINTER(Config)],
    "j2k_model_timing": [P, I, P],
    "j2k_model_autotune": [P, ctypes.c_longlong, ctypes.c_char_p, ctypes.c_size_t],
    "j2k_forward": [P, P, P, P],
    "j2k_step": [P, P, P, P],
    "j2k_forward_batch_dev": [P, I, P, P, P, I, P],
    "j2k_rollout": [P, P, P, I, I, I, P, ctypes.POINTER(D)],
    "j2k_rollout_dev": [P, P, P, I, I, I, P, P],
    "j2k_iterate_window": [P, P, P, I, I, I, I],
    "j2k_mass": [P, P, I, ctypes.POINTER(D)],
    "j2k_energy": [P, P, I, ctypes.POINTER(D)],
    "j2k_project_mass": [P, P, I, F],
    "j2k_project_energy": [P, P, I, F],
    "j2k_fill_blank": [P, P, I, I, I, P],
    "j2k_iterate": [P, P, P, P, I, I, I, I, F, ctypes.POINTER(I), ctypes.POINTER(F), ctypes.POINTER(D)],
    "j2k_iterate_dev": [P, P, P, P, I, I, I, I, F, ctypes.POINTER(I), ctypes.POINTER(F), ctypes.POINTER(D), P],
    "j2k_cell_forward_dev": [P, P, P, P, I, I, I, P, P],
    "j2k_cell_forward_dev_noalloc": [P, P, P, P, I, I, I, P, P],
    "j2k_server_create": [P, ctypes.c_size_t, I, PP],
    "j2k_server_submit": [P, P, P, P, I, I, I, ctypes.POINTER(ctypes.c_uint64)],
    "j2k_server_submit_rollout": [P, P, P, I, I, I, I, ctypes.POINTER(ctypes.c_uint64)],
    "j2k_server_windows": [P, ctypes.c_uint64, I, ctypes.POINTER(I), P, P, P],
    "j2k_server_poll": [P, ctypes.c_uint64, ctypes.POINTER(Result), P],
    "j2k_server_drain": [P],
    "j2k_server_stats": [P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(I),
                         ctypes.POINTER(ctypes.c_size_t)],
    "j2k_server_destroy": [P],
    "j2k_score_dev": [P, I, P, P, P, P, P, I, P, P],
    "j2k_score": [P, I, P, P, P, P, P, I, P],
    "j2k_correlation_dev": [P, F, F, P, F, F, I, P, P],
    "j2k_rollout_truth": [P, P, P, P, P, P, P, I, P],
    "j2k_rollout_truth_dev": [P, P, P, P, P, P, P, I, P, P],
    "j2k_excess_energy": [P, I, P, F, P],
    "j2k_excess_energy_dev": [P, I, P, F, P, P],
    "j2k_project_excess_energy": [P, I, P, P, F, F, F],
    "j2k_project_excess_energy_dev": [P, I, P, P, F, F, F, P],
    "j2k_sw_create": [ctypes.POINTER(SwParams), I, PP],
    "j2k_sw_destroy": [P],
    "j2k_sw_set_kernel": [P, I],
    "j2k_sw_steps": [P, P, P, P, I, I],
    "j2k_sw_steps_dev": [P, P, P, P, I, I, P],
    "j2k_sw_trajectories_dev": [P, P, P, P, I, P, I, P, P, P, P, P, P],
    "j2k_sw_trajectories": [P, P, P, P, I, P, P, I, P, P, I],
    "j2k_sw_drop_field": [ctypes.POINTER(SwParams), F, F, F, F, P],
}

_lib = None


def lib() -> ctypes.CDLL:
    """The loaded libj2k.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        # torch ships its own libamdhip64 under the same soname as the system
        # ROCm one libj2k links. Whichever loads first serves the whole
        # process, and torch fails to find a GPU when it is handed the other
        # runtime. Load torch first when it is installed, so both share its runtime.
        if "torch" not in sys.modules:
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
        if not LIB_PATH.exists():
            raise RuntimeError(f"libj2k.so is not built ({LIB_PATH}); run `python -c 'import __graft_entry__ as g; "
                               f"g.build()'` or `make -C jacobi2000_amd/csrc`")
        l = ctypes.CDLL(str(LIB_PATH))
        l.j2k_last_error.restype = ctypes.c_char_p
        l.j2k_last_error.argtypes = []
        for name, args in PROTOS.items():
            fn = getattr(l, name)
            fn.argtypes = args
            fn.restype = I
        _lib = l
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        raise J2KError(rc, lib().j2k_last_error().decode(errors="replace"))


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
