"""The C-ABI library on the CPU: it loads, exports every symbol include/j2k.h
declares, and its host-side parts (config/safetensors loader, blank fill)
behave like the reference's loader (crates/engine/src/{config,weights}.rs).
No compute call is made without a GPU."""

import ctypes
import json
import re
from pathlib import Path

import numpy as np
import pytest

import jacobi2000_amd as j
from jacobi2000_amd import _lib
from jacobi2000_amd import checkpoint as ck
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "j2k.h").read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(j2k_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = j.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.PROTOS) | {"j2k_last_error"}


def _create(st: bytes, cfg: str):
    buf = (ctypes.c_uint8 * len(st)).from_buffer_copy(st)
    h = ctypes.c_void_p()
    rc = j.lib().j2k_model_create(buf, len(st), cfg.encode(), 0, 0, ctypes.byref(h))
    return rc, j.lib().j2k_last_error().decode()


@pytest.fixture(scope="module")
def tiny_bytes(fixture_dir):
    return (fixture_dir / "model.safetensors").read_bytes(), (fixture_dir / "config.json").read_text()


def test_invalid_json_is_bad_config(tiny_bytes):
    rc, msg = _create(tiny_bytes[0], "{not json")
    assert rc == _lib.ERR_BAD_CONFIG and "invalid JSON" in msg


def test_missing_field_is_bad_config(tiny_bytes):
    cfg = json.loads(tiny_bytes[1])
    del cfg["dt"]
    rc, msg = _create(tiny_bytes[0], json.dumps(cfg))
    assert rc == _lib.ERR_BAD_CONFIG and "`dt`" in msg


def test_missing_tensor_names_it(tiny_bytes):
    t = O.read_safetensors(tiny_bytes[0])
    del t["blocks.0.ff1.weight"]
    rc, msg = _create(ck.write_safetensors(t), tiny_bytes[1])
    assert rc == _lib.ERR_MISSING_TENSOR and "blocks.0.ff1.weight" in msg


def test_non_f32_tensor_is_bad_dtype(tiny_bytes):
    st = pytest.importorskip("safetensors.numpy")
    t = O.read_safetensors(tiny_bytes[0])
    t = {k: (v.astype(np.float16) if k == "pos" else v) for k, v in t.items()}
    rc, msg = _create(st.save(t), tiny_bytes[1])
    assert rc == _lib.ERR_BAD_DTYPE and "pos" in msg


def test_truncated_buffer_is_shape_error(tiny_bytes):
    rc, _ = _create(tiny_bytes[0][:1000], tiny_bytes[1])
    assert rc == _lib.ERR_SHAPE


def test_unknown_keys_are_ignored_and_mask_defaults(tiny_bytes):
    """config.rs reads a fixed field set; `mask` defaults to Full and
    energy_r_max to 1.0 (config.rs:33-39). Past the parser the CPU-only
    container reports NoAdapter, the reference's `Error::NoAdapter`."""
    cfg = json.loads(tiny_bytes[1])
    cfg.pop("mask")
    cfg["excess_r"] = 3.0
    rc, msg = _create(tiny_bytes[0], json.dumps(cfg))
    if j.device_count() == 0:
        assert rc == _lib.ERR_NO_ADAPTER, msg
    else:
        assert rc == _lib.OK, msg


def test_fill_blank_matches_oracle():
    rng = np.random.default_rng(0)
    for kind, T in (("forward", 4), ("rewind", 5), ("reconstruction", 3)):
        g = rng.normal(0, 1, (3, T, 6, 7)).astype(np.float32)
        k = ck.mask_of(kind, T, 6, 7, seed=2, frac=0.1)
        assert np.array_equal(j.fill_blank(g, k), O.fill_blank(g, k))
    k = np.zeros((2, 3, 3), np.uint8)
    g = rng.normal(0, 1, (3, 2, 3, 3)).astype(np.float32)
    assert np.array_equal(j.fill_blank(g, k), O.fill_blank(g, k))


def test_null_arguments_are_invalid():
    rc = j.lib().j2k_forward(None, None, None, None)
    assert rc == _lib.ERR_INVALID
    assert "NULL" in j.lib().j2k_last_error().decode()


def _raw_safetensors(entry: dict, data_bytes: int = 4096) -> bytes:
    import json
    import struct

    h = json.dumps({"embed.weight": entry}).encode()
    h += b" " * ((8 - len(h) % 8) % 8)
    return struct.pack("<Q", len(h)) + h + bytes(data_bytes)


@pytest.mark.parametrize("entry", [
    # offsets near 2^64: the end offset wraps when added to the data start
    {"dtype": "F32", "shape": [512], "data_offsets": [2 ** 64 - 4096, 2 ** 64 - 2048]},
    {"dtype": "F32", "shape": [512], "data_offsets": [-8, 2040]},       # negative offset
    {"dtype": "F32", "shape": [512], "data_offsets": [0.5, 2048.5]},    # fractional offsets
    {"dtype": "F32", "shape": [-1, -512], "data_offsets": [0, 2048]},   # negative dimensions
    {"dtype": "F32", "shape": [2 ** 40, 2 ** 40], "data_offsets": [0, 0]},  # element count overflows
    {"dtype": "F32", "shape": [1024], "data_offsets": [2048, 6144]},    # past the end of the buffer
])
def test_malformed_offsets_and_shapes_are_rejected(tiny_bytes, entry):
    """The reader rejects offsets or dimensions that are not non-negative
    integers inside the buffer (weights.rs reads them as u64 and bounds-checks
    the slice), without reading outside it."""
    rc, msg = _create(_raw_safetensors(entry), tiny_bytes[1])
    print(rc, msg)
    assert rc == _lib.ERR_SHAPE and "embed.weight" in msg
