// Fused MLP of a block (cpu.rs:308-312): X += ff2(GELU(ff1(H))) and, when
// a next block follows, H = LN(X) with its norm1 — one kernel, the ff-wide
// intermediate never leaves the CU.
//
// A block owns 64 tokens and all dim (= 128) output columns; its four waves
// split the columns (32 each). The block's H rows are staged in LDS once. The
// intermediate is produced 128 columns at a time: ff1 + bias + GELU into a
// 64 x 128 LDS tile, which the ff2 accumulators consume right away as their
// next K chunk. Weights (W1t [ff, dim], W2t [dim, ff], bf16, K-contiguous)
// stream from L2 as 16-byte MFMA fragments, a chunk ahead of their use. v_mfma_f32_16x16x32_bf16, f32
// accumulate. SPLIT: H is f32 and both the staged H and the intermediate are
// held as bf16 hi + lo pairs (two MFMAs per tile), as in the split GEMM.
#include "common.h"
#include "kernels.h"

namespace j2k {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

constexpr int MBM = 64, MD = 128, MLD = MD + 8;

__device__ __forceinline__ bf8v frag(const uint16_t* p) { return __builtin_bit_cast(bf8v, *reinterpret_cast<const uint4*>(p)); }

template <bool SPLIT, bool LN>
__global__ __launch_bounds__(256) void k_mlp_bf16(int M, int ff, const void* __restrict__ Hin,
                                                  const uint16_t* __restrict__ W1t, const float* __restrict__ b1,
                                                  const uint16_t* __restrict__ W2t, const float* __restrict__ b2,
                                                  float* __restrict__ X, const float* __restrict__ lng,
                                                  const float* __restrict__ lnb, void* __restrict__ Hout) {
    __shared__ __align__(16) uint16_t Hh[MBM * MLD];
    __shared__ __align__(16) uint16_t Hl[SPLIT ? MBM * MLD : 8];
    __shared__ __align__(16) uint16_t Fh[MBM * MLD];
    __shared__ __align__(16) uint16_t Fl[SPLIT ? MBM * MLD : 8];
    __shared__ float red[4][MBM];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * MBM, wn = wave * 32;
    const int r16 = lane & 15, kq = (lane >> 4) * 8, rq = (lane >> 4) * 4;

    // stage the block's H rows
    if (!SPLIT) {
        const uint16_t* A = static_cast<const uint16_t*>(Hin);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int v = tid + 256 * q, row = v >> 4, kv = v & 15, m = m0 + row;
            const uint4 val = m < M ? *reinterpret_cast<const uint4*>(A + (long)m * MD + kv * 8) : make_uint4(0, 0, 0, 0);
            *reinterpret_cast<uint4*>(&Hh[row * MLD + kv * 8]) = val;
        }
    } else {
        const float* A = static_cast<const float*>(Hin);
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int v = tid + 256 * q, row = v >> 5, kv = v & 31, m = m0 + row;
            const float4 f = m < M ? *reinterpret_cast<const float4*>(A + (long)m * MD + kv * 4) : make_float4(0.f, 0.f, 0.f, 0.f);
            const float e[4] = {f.x, f.y, f.z, f.w};
            uint32_t hi[4], lo[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                hi[i] = f2bf(e[i]);
                lo[i] = f2bf(e[i] - bf2f((uint16_t)hi[i]));
            }
            *reinterpret_cast<uint2*>(&Hh[row * MLD + kv * 4]) = make_uint2(hi[0] | (hi[1] << 16), hi[2] | (hi[3] << 16));
            *reinterpret_cast<uint2*>(&Hl[row * MLD + kv * 4]) = make_uint2(lo[0] | (lo[1] << 16), lo[2] | (lo[3] << 16));
        }
    }
    __syncthreads();

    f4v acc2[4][2];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc2[i][j] = f4v{0.f, 0.f, 0.f, 0.f};

    // Weight fragments of a whole 128-wide chunk are loaded one phase ahead:
    // W2 of chunk c while ff1 of chunk c runs, W1 of chunk c+1 while GELU and
    // ff2 of chunk c run, so their L2 latency hides behind the MFMAs.
    bf8v w1[4][2], w2[4][2];
#define J2K_LOAD_W1(c0_)                                                                                        \
    _Pragma("unroll") for (int k4 = 0; k4 < 4; k4++) _Pragma("unroll") for (int j = 0; j < 2; j++)          \
        w1[k4][j] = frag(W1t + (long)((c0_) + wn + j * 16 + r16) * MD + k4 * 32 + kq);
#define J2K_LOAD_W2(c0_)                                                                                        \
    _Pragma("unroll") for (int k4 = 0; k4 < 4; k4++) _Pragma("unroll") for (int j = 0; j < 2; j++)          \
        w2[k4][j] = frag(W2t + (long)(wn + j * 16 + r16) * ff + (c0_) + k4 * 32 + kq);
    J2K_LOAD_W1(0)
    for (int c0 = 0; c0 < ff; c0 += MD) {
        J2K_LOAD_W2(c0)
        // ff1 for intermediate columns c0 + wn .. + 32 of this wave
        f4v acc1[4][2];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc1[i][j] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k4 = 0; k4 < 4; k4++) {
            const int ks = k4 * 32;
            bf8v a[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = frag(&Hh[(i * 16 + r16) * MLD + ks + kq]);
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc1[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], w1[k4][j], acc1[i][j], 0, 0, 0);
            if (SPLIT) {
#pragma unroll
                for (int i = 0; i < 4; i++) a[i] = frag(&Hl[(i * 16 + r16) * MLD + ks + kq]);
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++) acc1[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], w1[k4][j], acc1[i][j], 0, 0, 0);
            }
        }
        float bb[2];
#pragma unroll
        for (int j = 0; j < 2; j++) bb[j] = b1[c0 + wn + j * 16 + r16];
        if (c0 + MD < ff) {
            J2K_LOAD_W1(c0 + MD)
        }
        __syncthreads();  // every wave has finished reading the previous intermediate chunk
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float f = gelu(acc1[i][j][r] + bb[j]);
                    const int o = (i * 16 + rq + r) * MLD + wn + j * 16 + r16;
                    const uint16_t h = f2bf(f);
                    Fh[o] = h;
                    if (SPLIT) Fl[o] = f2bf(f - bf2f(h));
                }
        __syncthreads();
        // ff2: output columns wn .. + 32, K chunk c0 .. c0 + 128
#pragma unroll
        for (int k4 = 0; k4 < 4; k4++) {
            const int ks = k4 * 32;
            bf8v a[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = frag(&Fh[(i * 16 + r16) * MLD + ks + kq]);
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc2[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], w2[k4][j], acc2[i][j], 0, 0, 0);
            if (SPLIT) {
#pragma unroll
                for (int i = 0; i < 4; i++) a[i] = frag(&Fl[(i * 16 + r16) * MLD + ks + kq]);
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++) acc2[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], w2[k4][j], acc2[i][j], 0, 0, 0);
            }
        }
    }
#undef J2K_LOAD_W1
#undef J2K_LOAD_W2

    // residual
    float bo[2];
#pragma unroll
    for (int j = 0; j < 2; j++) bo[j] = b2[wn + j * 16 + r16];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = m0 + i * 16 + rq + r;
            if (m >= M) continue;
#pragma unroll
            for (int j = 0; j < 2; j++) {
                float* p = X + (long)m * MD + wn + j * 16 + r16;
                const float v = *p + acc2[i][j][r] + bo[j];
                *p = v;
                acc2[i][j][r] = v;
            }
        }
    if (!LN) return;
    // LayerNorm of the updated rows (cpu.rs:49-65); a row spans 4 waves x 2 tiles x 16 lanes
    float mean[4][4], inv[4][4];
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                float s = 0.f;
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const float v = acc2[i][j][r];
                    s += pass == 0 ? v : (v - mean[i][r]) * (v - mean[i][r]);
                }
                s += __shfl_xor(s, 1, 64);
                s += __shfl_xor(s, 2, 64);
                s += __shfl_xor(s, 4, 64);
                s += __shfl_xor(s, 8, 64);
                if (r16 == 0) red[wave][i * 16 + rq + r] = s;
            }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = i * 16 + rq + r;
                const float tot = red[0][row] + red[1][row] + red[2][row] + red[3][row];
                if (pass == 0) mean[i][r] = tot / (float)MD;
                else inv[i][r] = 1.0f / sqrtf(tot / (float)MD + 1e-5f);
            }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int n = wn + j * 16 + r16;
        const float g = lng[n], b = lnb[n];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + i * 16 + rq + r;
                if (m >= M) continue;
                const float h = (acc2[i][j][r] - mean[i][r]) * inv[i][r] * g + b;
                if (SPLIT) static_cast<float*>(Hout)[(long)m * MD + n] = h;
                else static_cast<uint16_t*>(Hout)[(long)m * MD + n] = f2bf(h);
            }
    }
}

bool mlp_fused_ok(int dim, int ff) { return dim == MD && ff % MD == 0; }

void mlp_fused(hipStream_t s, int M, int dim, int ff, bool split, const void* H, const uint16_t* W1t, const float* b1,
               const uint16_t* W2t, const float* b2, float* X, const float* ln_g, const float* ln_b, void* Hout) {
    if (M <= 0) return;
    if (!mlp_fused_ok(dim, ff)) fail(J2K_ERR_LIMIT, "fused MLP needs dim == 128 and ff a multiple of 128");
    dim3 grid(cdiv(M, MBM)), block(256);
    const bool ln = ln_g != nullptr;
    if (split) {
        if (ln) hipLaunchKernelGGL((k_mlp_bf16<true, true>), grid, block, 0, s, M, ff, H, W1t, b1, W2t, b2, X, ln_g, ln_b, Hout);
        else hipLaunchKernelGGL((k_mlp_bf16<true, false>), grid, block, 0, s, M, ff, H, W1t, b1, W2t, b2, X, ln_g, ln_b, Hout);
    } else {
        if (ln) hipLaunchKernelGGL((k_mlp_bf16<false, true>), grid, block, 0, s, M, ff, H, W1t, b1, W2t, b2, X, ln_g, ln_b, Hout);
        else hipLaunchKernelGGL((k_mlp_bf16<false, false>), grid, block, 0, s, M, ff, H, W1t, b1, W2t, b2, X, ln_g, ln_b, Hout);
    }
    LAUNCHCHK("mlp_fused");
}

}  // namespace j2k
