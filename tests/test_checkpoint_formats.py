"""The checkpoint format extension (BF16 and F8_E4M3 block weights with a
per-tensor scale): the encoders against the OCP FP8 E4M3FN definition and
round-to-nearest-even bf16, the converter's round trip, and the C loader's
errors for malformed F8 tensors (parsing runs before any device is needed)."""

import ctypes

import numpy as np
import pytest

from jacobi2000_amd import _lib
from jacobi2000_amd import checkpoint as ck


def test_e4m3_known_codes():
    # OCP 8-bit floating point specification, E4M3: bias 7, max normal
    # 1.75 * 2^8 = 448 (0x7E), min normal 2^-6 (0x08), max subnormal
    # 0.875 * 2^-6 (0x07), min subnormal 2^-9 (0x01), NaN S.1111.111
    vals = [448.0, 1.0, 2.0 ** -6, 0.875 * 2.0 ** -6, 2.0 ** -9, 0.0, -1.5, 3.5, 240.0]
    codes = [0x7E, 0x38, 0x08, 0x07, 0x01, 0x00, 0xBC, 0x46, 0x77]
    assert list(ck.e4m3_encode(np.array(vals))) == codes
    assert ck.e4m3_encode(np.array([np.nan]))[0] == 0x7F
    assert list(ck.e4m3_encode(np.array([1e9, -600.0]))) == [0x7E, 0xFE]  # saturation


def test_e4m3_decode_matches_integer_form():
    # an independent statement of the same table: (8 + m) * 2^(e - 10) for
    # e > 0, m * 2^-9 for e == 0
    for c in range(256):
        e, m = (c >> 3) & 15, c & 7
        got = float(ck.e4m3_decode(np.array([c], np.uint8))[0])
        if e == 15 and m == 7:
            assert np.isnan(got)
            continue
        want = (m * 2.0 ** -9) if e == 0 else (8 + m) * 2.0 ** (e - 10)
        assert got == (-want if c & 0x80 else want), hex(c)


def test_e4m3_rounds_to_nearest_even():
    # halfway cases go to the even code; everything else to the nearest value
    assert list(ck.e4m3_encode(np.array([1.0625, 1.1875, 2.0 ** -10, 3 * 2.0 ** -10]))) == [0x38, 0x3A, 0x00, 0x02]
    rng = np.random.default_rng(0)
    x = rng.normal(0, 20, 20000)
    q = ck.e4m3_decode(ck.e4m3_encode(x)).astype(np.float64)
    grid = ck.e4m3_decode(np.arange(0x7F, dtype=np.uint8)).astype(np.float64)
    nearest = np.abs(np.abs(x)[:, None] - grid[None, :]).min(1)
    assert np.all(np.sign(q) * np.sign(x) >= 0)
    assert np.all(np.abs(np.abs(x) - np.abs(q)) <= nearest + 1e-12)
    codes = np.arange(256, dtype=np.uint8)
    v = ck.e4m3_decode(codes)
    ok = ~np.isnan(v)
    assert np.array_equal(ck.e4m3_encode(v[ok]), codes[ok])


def test_bf16_rounding():
    x = np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 1.0 + 2.0 ** -9], np.float32)
    bits = ck.f32_to_bf16_bits(x)
    back = (bits.astype(np.uint32) << 16).view(np.float32)
    assert list(back) == [1.0, 1.0 + 2.0 ** -6, -2.5, 1.0]


@pytest.mark.parametrize("dtype", ["BF16", "F8_E4M3"])
def test_convert_round_trip(dtype):
    st, cj = ck.random_cell_st(dim=32, layers=2, heads=2, ff=64, seed=5, std=0.1)
    src = ck.read_safetensors(st)
    conv = ck.convert_checkpoint(st, dtype)
    got = ck.read_safetensors(conv)
    assert set(got) == set(src)
    for name, a in src.items():
        b = got[name]
        if name.endswith("q.weight") or name.endswith("ff2.weight"):
            rel = np.abs(b - a).max() / np.abs(a).max()
            bound = 2.0 ** -8 if dtype == "BF16" else 2.0 ** -4
            assert 0 < rel <= bound, (name, rel)
        elif not any(name.endswith(f"{n}.weight") for n in ck.BLOCK_LINEARS):
            assert np.array_equal(a, b), name


def _create(st):
    h = ctypes.c_void_p()
    buf = (ctypes.c_uint8 * len(st)).from_buffer_copy(st)
    _, cj = ck.random_cell_st(dim=32, layers=2, heads=2, ff=64, seed=5)
    rc = _lib.lib().j2k_model_create(buf, len(st), cj.encode(), 0, _lib.PREC_FP8, ctypes.byref(h))
    msg = _lib.lib().j2k_last_error().decode()
    if rc == _lib.OK:
        _lib.lib().j2k_model_destroy(h)
    return rc, msg


def test_loader_f8_needs_its_scale():
    st, _ = ck.random_cell_st(dim=32, layers=2, heads=2, ff=64, seed=5)
    t = ck.read_safetensors(ck.convert_checkpoint(st, "F8_E4M3"))
    codes = {n: ck.e4m3_encode(a) for n, a in t.items() if n.endswith("blocks.0.q.weight")}
    rest = {n: a for n, a in t.items() if n not in codes}
    rc, msg = _create(ck.write_safetensors({**rest, **codes}, {n: "F8_E4M3" for n in codes}))
    assert rc == _lib.ERR_MISSING_TENSOR and "blocks.0.q.weight.scale" in msg


def test_loader_reads_converted_checkpoints():
    st, _ = ck.random_cell_st(dim=32, layers=2, heads=2, ff=64, seed=5)
    for dtype in ("BF16", "F8_E4M3"):
        rc, msg = _create(ck.convert_checkpoint(st, dtype))
        # parsing succeeds; without a GPU the device check is what fails
        assert rc in (_lib.OK, _lib.ERR_NO_ADAPTER), msg


def test_fp8_restatement_unquantized_equals_oracle():
    """The numpy FP8 restatement used by the GPU parity test reduces to the
    oracle's f64 forward without quantization, and quantization moves it by
    a few percent (E4M3 keeps 3 mantissa bits)."""
    from oracle import oracle as O
    from tests import fp8_emulation as E

    st, cj = ck.random_cell_st(dim=32, layers=2, heads=2, ff=64, seed=6, std=0.1)
    m = O.FormC.load(st, cj)
    T, ny, nx = 3, 5, 4
    g = ck.drops_grid(T, ny, nx, 2, 7)
    k = ck.mask_of("rewind", T, ny, nx)
    x0 = O.fill_blank(g, k)
    meta = ck.meta_of(1.0, 0.09, 9.81, 1.0, 1e-3)
    ref = m.forward(x0, k, meta, f64=True)
    exact = E.cell_forward(m.w, m.cfg, x0, k, meta, fp8=False)
    assert np.abs(exact - ref).max() < 1e-10
    q = E.cell_forward(m.w, m.cfg, x0, k, meta, fp8=True)
    rel = np.linalg.norm(q - exact) / np.linalg.norm(exact)
    assert 1e-3 < rel < 0.1
