"""jacobi2000 engine forward pass on MI355X (gfx950): hand-written HIP
kernels in libj2k.so behind the C ABI of include/j2k.h, with a Python mirror
of the reference engine API (crates/engine/src/lib.rs)."""

from ._lib import J2KError, lib  # noqa: F401
from .engine import GpuModel, Model, device_count, fill_blank  # noqa: F401
