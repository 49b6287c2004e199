// The cell_st embedding of one token and its LayerNorm, for the kernels
// that take a token with L = D / 8 lanes (lane j: columns 8 j .. 8 j + 7):
// the embed linear (cpu.rs:26-45 for one row: bias first, then the four
// features in order, zero features skipped) plus the meta vector, then
// LN1 (cpu.rs:49-65) with the row sums grouped over the token's lanes.
// One definition, so every kernel that embeds gives the same bits.
#pragma once

#include "common.h"

namespace j2k {

// Sum over the L lanes of one token's group (L a power of two up to 32).
template <int L>
__device__ __forceinline__ float group_sum(float x) {
    static_assert(L >= 4 && L <= 32 && (L & (L - 1)) == 0, "group of 4 to 32 lanes");
    x += quad_xor1(x);
    x += quad_xor2(x);
    if (L > 4) x += swz_xor<4>(x);
    if (L > 8) x += swz_xor<8>(x);
    if (L > 16) x += swz_xor<16>(x);
    return x;
}

typedef float emb_f4v __attribute__((ext_vector_type(4)));

// a: the token's columns d0 .. d0 + 7 of the embedding (X), o: of LN1(X)
template <int D>
__device__ __forceinline__ void embed_ln_token(const float4 f, const float* __restrict__ we,
                                               const float* __restrict__ be, const float* __restrict__ mv,
                                               const float* __restrict__ g, const float* __restrict__ b, int d0,
                                               float (&a)[8], float (&o)[8]) {
    constexpr int L = D / 8;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const emb_f4v bv = *reinterpret_cast<const emb_f4v*>(be + d0 + 4 * h);
        const emb_f4v w0 = *reinterpret_cast<const emb_f4v*>(we + d0 + 4 * h);
        const emb_f4v w1 = *reinterpret_cast<const emb_f4v*>(we + D + d0 + 4 * h);
        const emb_f4v w2 = *reinterpret_cast<const emb_f4v*>(we + 2 * D + d0 + 4 * h);
        const emb_f4v w3 = *reinterpret_cast<const emb_f4v*>(we + 3 * D + d0 + 4 * h);
        const emb_f4v mw = *reinterpret_cast<const emb_f4v*>(mv + d0 + 4 * h);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float v = bv[k];
            if (f.x != 0.f) v += f.x * w0[k];
            if (f.y != 0.f) v += f.y * w1[k];
            if (f.z != 0.f) v += f.z * w2[k];
            if (f.w != 0.f) v += f.w * w3[k];
            v += mw[k];
            a[4 * h + k] = v;
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) s += a[k];
    const float mean = group_sum<L>(s) / (float)D;
    float q = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float c = a[k] - mean;
        q += c * c;
    }
    const float inv = 1.0f / sqrtf(group_sum<L>(q) / (float)D + 1e-5f);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const emb_f4v gv = *reinterpret_cast<const emb_f4v*>(g + d0 + 4 * h);
        const emb_f4v bb = *reinterpret_cast<const emb_f4v*>(b + d0 + 4 * h);
#pragma unroll
        for (int k = 0; k < 4; k++) o[4 * h + k] = (a[4 * h + k] - mean) * inv * gv[k] + bb[k];
    }
}

}  // namespace j2k
