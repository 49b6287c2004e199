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

// x of lane ^ 1 and of lane ^ 2 (DPP quad permutations: a VALU move, where
// __shfl_xor goes through the LDS unit's permute).
__device__ __forceinline__ float quad_xor1(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0xB1, 0xF, 0xF, true));
}
__device__ __forceinline__ float quad_xor2(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x4E, 0xF, 0xF, true));
}
// x of lane j of the quad, to all four (j a constant once unrolled).
__device__ __forceinline__ float quad_bcast(float x, int j) {
    const int v = __builtin_bit_cast(int, x);
    switch (j & 3) {
        case 0: return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(v, 0x00, 0xF, 0xF, true));
        case 1: return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(v, 0x55, 0xF, 0xF, true));
        case 2: return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(v, 0xAA, 0xF, 0xF, true));
        default: return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(v, 0xFF, 0xF, 0xF, true));
    }
}
// Reduce-scatter over a quad: lane j (= sub) gets p[j] summed over the four
// lanes as (lane 0 + lane 1) + (lane 2 + lane 3), the value and rounding of
// the butterfly x += quad_xor1(x); x += quad_xor2(x) on p[j], with three
// exchanges for the four sums instead of eight.
__device__ __forceinline__ float quad_scatter4(const float (&p)[4], int sub) {
    const bool odd = sub & 1, hi = sub & 2;
    const float a = (odd ? p[1] : p[0]) + quad_xor1(odd ? p[0] : p[1]);  // pair sums of keys 0 / 1
    const float b = (odd ? p[3] : p[2]) + quad_xor1(odd ? p[2] : p[3]);  // and of keys 2 / 3
    return (hi ? b : a) + quad_xor2(hi ? a : b);
}

// x of lane ^ M within each 32 lanes, M = 4, 8 or 16 (ds_swizzle in its
// bit-mask mode: the LDS unit's crossbar without an address or a bank access).
template <int M>
__device__ __forceinline__ float swz_xor(float x) {
    static_assert(M == 4 || M == 8 || M == 16, "swizzle xor within 32 lanes");
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, x), 0x1F | (M << 10)));
}

// x of lane ^ o for o = 1 .. 16 (DPP or swizzle), or 32 (a permute).
__device__ __forceinline__ float lane_xor(float x, int o) {
    switch (o) {
        case 1: return quad_xor1(x);
        case 2: return quad_xor2(x);
        case 4: return swz_xor<4>(x);
        case 8: return swz_xor<8>(x);
        case 16: return swz_xor<16>(x);
        default: return __shfl_xor(x, o, WAVE);
    }
}
__device__ __forceinline__ double lane_xor(double x, int o) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const float lo = lane_xor(__uint_as_float((unsigned)b), o), hi = lane_xor(__uint_as_float((unsigned)(b >> 32)), o);
    return __builtin_bit_cast(double, (unsigned long long)__float_as_uint(lo) |
                                          ((unsigned long long)__float_as_uint(hi) << 32));
}

// Butterfly over the wave, lanes xor 32, 16, ..., 1: every lane ends with
// the total.
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += lane_xor(v, o);
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, lane_xor(v, o));
    return v;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += lane_xor(v, o);
    return v;
}

// Exact-erf GELU, crates/engine/src/cpu.rs:67-74 (device erff from ROCm's math library).
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

// GELU of the served forms, whose activations are rounded to bf16 (pairs)
// right after: erfc(z) = t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-z^2),
// t = 1 / (1 + p z), z = |x| / sqrt(2) (Abramowitz & Stegun 7.1.26, Hastings'
// approximation, |error of erf| <= 1.5e-7). With h = |x| erfc(z) / 2,
// GELU(x) = max(x, 0) - h on both sides (x - h for x >= 0, and for x < 0
// x erfc(z) / 2 = -h, with no cancellation). The 1/2 and 1/sqrt2 ride in the
// constants, exp(-z^2) = 2^(x^2 (-log2(e) / 2)): twelve VALU operations and
// two transcendentals.
__device__ __forceinline__ float gelu_fast(float x) {
    const float a = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f * 0.70710678118654752f, a, 1.0f));
    float p = fmaf(0.5f * 1.061405429f, t, 0.5f * -1.453152027f);
    p = fmaf(p, t, 0.5f * 1.421413741f);
    p = fmaf(p, t, 0.5f * -0.284496736f);
    p = fmaf(p, t, 0.5f * 0.254829592f);
    const float h = p * t * __builtin_amdgcn_exp2f(x * x * -0.72134752044448170f);
    return fmaf(-a, h, fmaxf(x, 0.0f));
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

// Output o of the meta linear (cpu.rs:26-45 for one row: bias first, then
// the inputs in order).
__device__ __forceinline__ float meta_out(int md, int dim, const float* meta, const float* w, const float* b, int o) {
    float acc = b[o];
    for (int i = 0; i < md; i++) acc += meta[i] * w[i * dim + o];
    return acc;
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// Workgroups are dealt round-robin over the 8 XCDs (block b runs on XCD
// b mod 8). This bijection on [0, G) hands each XCD one contiguous run of
// work items, so neighbouring items share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int G) {
    const int q = G >> 3, r = G & 7, k = b & 7;
    return k * q + min(k, r) + (b >> 3);
}

}  // namespace j2k
