// Checkpoint loading: a small JSON reader, the safetensors reader and the
// config.json schema (host only).
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/j2k.h"
#include "error.h"

namespace j2k {

// A parsed JSON value (null, bool, number, string, array, object).
struct Json {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    static Json parse(const std::string& text);  // throws Error(J2K_ERR_BAD_CONFIG)
    const Json* find(const std::string& key) const;
};

struct Tensor {
    std::vector<int64_t> shape;
    std::vector<float> data;      // f32 values (F8 tensors: decoded times their scale)
    std::vector<uint8_t> f8;      // F8_E4M3 tensors: the codes as stored
    float f8_scale = 0.f;         // and their dequantization scale (`<name>.scale`)
    int64_t numel() const {
        int64_t n = 1;
        for (auto s : shape) n *= s;
        return n;
    }
};

using Weights = std::map<std::string, Tensor>;

// crates/engine/src/weights.rs:29-48: 8-byte LE header length, JSON header
// {name: {dtype, shape, data_offsets}}, LE data. F32 as the reference; the
// extension of the format adds BF16 (widened exactly) and F8_E4M3 (OCP
// E4M3FN codes) whose f32 value is code * `<name>.scale`, an F32 tensor of
// one element stored beside it. Other dtypes fail with J2K_ERR_BAD_DTYPE
// (weights.rs:39 panics).
Weights read_safetensors(const uint8_t* bytes, size_t n);

// OCP FP8 E4M3FN: 1 sign, 4 exponent (bias 7), 3 mantissa bits, no
// infinities, S.1111.111 is NaN, largest finite 448. Encoding rounds to
// nearest even and saturates to +-448 (NaN stays NaN).
float e4m3_decode(uint8_t code);
uint8_t e4m3_encode(float v);

// weights.rs:52-54: a named tensor or J2K_ERR_MISSING_TENSOR naming it.
const Tensor& get(const Weights& w, const std::string& name);

// crates/engine/src/config.rs:70-100 (+ the cell_st keys). head_dim from
// head.to_cells.weight.shape[0] and ff_inner from blocks.0.ff1.weight.shape[1]
// as config.rs:72-73; cell_st has no CellHead so its head_dim is unused.
j2k_config parse_config(const std::string& json_text, const Weights& w, int precision);

}  // namespace j2k
