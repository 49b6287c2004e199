// Shared device helpers and host error plumbing for libj2k (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/j2k.h"
#include "error.h"

namespace j2k {

inline void hip_check(hipError_t e, const char* what, int code = J2K_ERR_POLL) {
    if (e != hipSuccess) fail(e == hipErrorOutOfMemory ? J2K_ERR_OOM : code, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(x) ::j2k::hip_check((x), #x)
#define LAUNCHCHK(name) ::j2k::hip_check(hipGetLastError(), name)

constexpr int WAVE = 64;

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, WAVE));
    return v;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// Exact-erf GELU, crates/engine/src/cpu.rs:67-74 (device erff from ROCm's math library).
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

// GELU of the served forms, whose activations are rounded to bf16 (pairs)
// right after: erfc(z) = t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-z^2),
// t = 1 / (1 + p z), z = |x| / sqrt(2) (Abramowitz & Stegun 7.1.26, Hastings'
// approximation, |error of erf| <= 1.5e-7), then 1 + erf(x/sqrt2) = 2 - erfc
// for x >= 0 and erfc for x < 0 (no cancellation on the negative side). About
// a third of the instructions of the erff form.
__device__ __forceinline__ float gelu_fast(float x) {
    const float z = fabsf(x) * 0.70710678118654752f;
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, z, 1.0f));
    float p = fmaf(1.061405429f, t, -1.453152027f);
    p = fmaf(p, t, 1.421413741f);
    p = fmaf(p, t, -0.284496736f);
    p = fmaf(p, t, 0.254829592f);
    const float q = p * t * __builtin_amdgcn_exp2f(-(z * z) * 1.4426950408889634f);
    return 0.5f * x * (x >= 0.f ? 2.0f - q : q);
}

// f32 -> bf16 rounding to nearest even, on gfx950's conversion instruction
// (v_cvt_pk_bf16_f32; NaN stays a quiet NaN); f2bf2 packs two (a low).
typedef __bf16 j2k_bf2v __attribute__((ext_vector_type(2)));
typedef float j2k_f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ uint32_t f2bf2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((j2k_f2v{a, b}), j2k_bf2v));
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// Workgroups are dealt round-robin over the 8 XCDs (block b runs on XCD
// b mod 8). This bijection on [0, G) hands each XCD one contiguous run of
// work items, so neighbouring items share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int G) {
    const int q = G >> 3, r = G & 7, k = b & 7;
    return k * q + min(k, r) + (b >> 3);
}

}  // namespace j2k
