// Stencil gather-attention (crates/engine/src/cpu.rs:97-164 restricted to the
// allowed keys, as attention.wgsl:47-107 does): per query and head,
// s = q.k / sqrt(hd) over the stencil's keys, softmax, o = sum p v.
//
// A group of hd/4 lanes per (token, head), each lane holding 4 of the head's
// dimensions: the dot products are finished by xor shuffles inside the
// group, every lane runs the softmax (the single-pass running-max form,
// rescaled on a new max, in f32) and weights its own 4 dimensions. A block
// covers 1024/dim consecutive tokens of one segment with all their heads, so
// a 256-token grid of the reference forward already fills 32 blocks and the
// stencil's overlapping key rows are served from L1/L2.
#include "common.h"
#include "kernels.h"

namespace j2k {

// Four consecutive elements as f32.
__device__ __forceinline__ void load4(const float* p, float* out) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
    out[3] = v.w;
}
__device__ __forceinline__ void load4(const uint16_t* p, float* out) {
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    out[0] = __uint_as_float(v.x << 16);
    out[1] = __uint_as_float(v.x & 0xffff0000u);
    out[2] = __uint_as_float(v.y << 16);
    out[3] = __uint_as_float(v.y & 0xffff0000u);
}

__device__ __forceinline__ void store_out(float* p, float v) { *p = v; }
__device__ __forceinline__ void store_out(uint16_t* p, float v) { *p = f2bf(v); }

// Calls op(key_local_index) for every allowed key of query `q` (local index
// within the segment), in the order: patch stencil ascending token index
// (mask.rs:18-30); full ascending; cell_st time, then y, then x offsets.
template <class Op>
__device__ __forceinline__ void for_keys(const AttnGeo& g, const Seg& sg, int q, Op&& op) {
    if (g.kind == ATTN_FULL) {
        const int n = sg.T * sg.ny * sg.nx;
        for (int k = 0; k < n; k++) op(k);
        return;
    }
    const int n = sg.ny * sg.nx;
    const int t = q / n, rem = q - t * n, y = rem / sg.nx, x = rem - y * sg.nx;
    if (g.kind == ATTN_PATCH_STENCIL) {
        const int r = g.radius * g.dilation, d = g.dilation;
        for (int dy = -r; dy <= r; dy += d) {
            const int ky = y + dy;
            if (ky < 0 || ky >= sg.ny) continue;
            for (int dx = -r; dx <= r; dx += d) {
                const int kx = x + dx;
                if (kx < 0 || kx >= sg.nx) continue;
                op(t * n + ky * sg.nx + kx);
            }
        }
        return;
    }
    const int rs = g.radius, rt = g.time_radius;
    for (int tt = t - rt; tt <= t + rt; tt++) {
        if (tt < 0 || tt >= sg.T) continue;
        for (int dy = -rs; dy <= rs; dy++) {
            int ky = y + dy;
            ky += ky < 0 ? sg.ny : (ky >= sg.ny ? -sg.ny : 0);
            for (int dx = -rs; dx <= rs; dx++) {
                int kx = x + dx;
                kx += kx < 0 ? sg.nx : (kx >= sg.nx ? -sg.nx : 0);
                op((tt * sg.ny + ky) * sg.nx + kx);
            }
        }
    }
}

template <int HD, class T>
__global__ __launch_bounds__(256) void k_attention(const T* __restrict__ qkv, T* __restrict__ out, int dim, int heads,
                                                   AttnGeo geo, const Seg* __restrict__ segs,
                                                   const Tile* __restrict__ tiles, int tpb) {
    constexpr int G = HD / 4;  // lanes of one (token, head)
    const Tile tl = tiles[blockIdx.x];
    const Seg sg = segs[tl.seg];
    const int grp = threadIdx.x / G, sub = threadIdx.x - grp * G;
    const int lt = grp / heads, head = grp - lt * heads;
    // a group past the tile's end follows the tile's first token (so that its
    // shuffles pair up) and stores nothing
    const bool live = lt < tpb && tl.start + lt < sg.T * sg.ny * sg.nx;
    const int q = live ? tl.start + lt : tl.start;
    const long row = (long)sg.base + q, ld = 3L * dim;
    float qv[4], acc[4] = {0.f, 0.f, 0.f, 0.f};
    load4(qkv + row * ld + head * HD + sub * 4, qv);
    const float scale = 1.0f / sqrtf((float)HD);
    float mx = -INFINITY, l = 0.f;
    const T* kbase = qkv + (long)sg.base * ld + dim + head * HD + sub * 4;
    for_keys(geo, sg, q, [&](int k) {
        float kv[4];
        load4(kbase + (long)k * ld, kv);
        float dot = 0.f;
#pragma unroll
        for (int e = 0; e < 4; e++) dot += qv[e] * kv[e];
#pragma unroll
        for (int off = 1; off < G; off <<= 1) dot += __shfl_xor(dot, off, G);
        const float s = dot * scale;
        const float mn = fmaxf(mx, s);
        const float corr = expf(mx - mn), p = expf(s - mn);
        l = l * corr + p;
        load4(kbase + dim + (long)k * ld, kv);
#pragma unroll
        for (int e = 0; e < 4; e++) acc[e] = acc[e] * corr + p * kv[e];
        mx = mn;
    });
    if (!live) return;
    const float inv = 1.0f / l;
    T* o = out + row * dim + head * HD + sub * 4;
#pragma unroll
    for (int e = 0; e < 4; e++) store_out(o + e, acc[e] * inv);
}

int attn_tokens_per_block(int dim) { return 1024 / dim; }

template <class T>
static void launch_attention(hipStream_t s, const T* qkv, T* out, int dim, int heads, AttnGeo geo, const Seg* segs,
                             const Tile* tiles, int n_tiles, int tpb) {
    if (n_tiles <= 0) return;
    const int hd = dim / heads;
    if (heads > 256 || dim % heads) fail(J2K_ERR_LIMIT, "attention: heads must divide dim and be <= 256");
    dim3 grid(n_tiles), block(256);
    switch (hd) {
        case 8: hipLaunchKernelGGL((k_attention<8, T>), grid, block, 0, s, qkv, out, dim, heads, geo, segs, tiles, tpb); break;
        case 16: hipLaunchKernelGGL((k_attention<16, T>), grid, block, 0, s, qkv, out, dim, heads, geo, segs, tiles, tpb); break;
        case 32: hipLaunchKernelGGL((k_attention<32, T>), grid, block, 0, s, qkv, out, dim, heads, geo, segs, tiles, tpb); break;
        case 64: hipLaunchKernelGGL((k_attention<64, T>), grid, block, 0, s, qkv, out, dim, heads, geo, segs, tiles, tpb); break;
        default: fail(J2K_ERR_LIMIT, "attention: head size must be 8, 16, 32 or 64");
    }
    LAUNCHCHK("attention");
}

void attention_f32(hipStream_t s, const float* qkv, float* out, int dim, int heads, AttnGeo geo, const Seg* segs,
                   const Tile* tiles, int n_tiles, int tpb) {
    launch_attention<float>(s, qkv, out, dim, heads, geo, segs, tiles, n_tiles, tpb);
}

void attention_bf16(hipStream_t s, const uint16_t* qkv, uint16_t* out, int dim, int heads, AttnGeo geo,
                    const Seg* segs, const Tile* tiles, int n_tiles, int tpb) {
    launch_attention<uint16_t>(s, qkv, out, dim, heads, geo, segs, tiles, n_tiles, tpb);
}

}  // namespace j2k
