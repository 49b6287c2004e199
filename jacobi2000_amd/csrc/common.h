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

// bf16 <-> f32, round to nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
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
