// Row-panel bf16 GEMM for the trunk's tall-skinny products (M = tokens,
// N, K <= a few hundred) on v_mfma_f32_16x16x32_bf16, f32 accumulate.
//
// A block owns 64 rows and 64*CT columns (CT = N/64 when N <= 512, so one
// block covers whole rows). Its four waves split the columns; each wave holds
// 4 x CT MFMA tiles of 16 x 16. The A panel is staged into LDS 128 K-columns
// at a time (rows padded by 8 bf16 so the 16-row fragment reads hit distinct
// banks) and read once from HBM per column block; the weights (Bt [N, K],
// K-contiguous) stream from L2 in 16-byte fragments.
//
// Epilogues: bf16 store, GELU + bf16 store, residual add into the f32
// stream, and residual add followed by the LayerNorm of the updated rows
// (needs the block to own whole rows), which writes the next GEMM's bf16
// input so no separate LayerNorm pass reads the stream again.
#include "common.h"
#include "kernels.h"

namespace j2k {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

constexpr int PBM = 64, PKC = 128, PLDA = PKC + 8;

// SPLIT: A is f32, staged as a bf16 pair hi = bf16(a), lo = bf16(a - hi),
// and every tile takes two MFMAs (hi and lo against the same bf16 weight
// fragment): the activations keep about 16 mantissa bits, so the result
// carries only the weights' bf16 rounding, which is the same on every call.
// Outputs (C, H) are then f32.
template <int CT, int EPI, bool SPLIT>
__global__ __launch_bounds__(256) void k_panel_bf16(int M, int N, int K, const void* __restrict__ Av,
                                                    const uint16_t* __restrict__ Bt, const float* __restrict__ bias,
                                                    void* __restrict__ Cv, float* __restrict__ X,
                                                    const float* __restrict__ lng, const float* __restrict__ lnb,
                                                    void* __restrict__ Hv) {
    __shared__ __align__(16) uint16_t As[PBM * PLDA];
    __shared__ __align__(16) uint16_t Al[SPLIT ? PBM * PLDA : 8];
    __shared__ float red[4][PBM];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * PBM;
    const int nw = blockIdx.y * (64 * CT) + wave * (16 * CT);
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    f4v acc[4][CT];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < CT; j++) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};

    for (int kc = 0; kc < K; kc += PKC) {
        const int kw = min(PKC, K - kc);  // a multiple of 32
        const int vpr = kw / 8;           // 16-byte vectors per row in this chunk
        __syncthreads();
        if (!SPLIT) {
            const uint16_t* A = static_cast<const uint16_t*>(Av);
            for (int c = tid; c < PBM * vpr; c += 256) {
                const int row = c / vpr, v = c - row * vpr, m = m0 + row;
                uint4 val = m < M ? *reinterpret_cast<const uint4*>(A + (long)m * K + kc + v * 8) : make_uint4(0, 0, 0, 0);
                *reinterpret_cast<uint4*>(&As[row * PLDA + v * 8]) = val;
            }
        } else {
            const float* A = static_cast<const float*>(Av);
            const int fpr = kw / 4;  // float4 per row
            for (int c = tid; c < PBM * fpr; c += 256) {
                const int row = c / fpr, v = c - row * fpr, m = m0 + row;
                float4 f = m < M ? *reinterpret_cast<const float4*>(A + (long)m * K + kc + v * 4) : make_float4(0.f, 0.f, 0.f, 0.f);
                const float e[4] = {f.x, f.y, f.z, f.w};
                uint16_t hi[4], lo[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    hi[q] = f2bf(e[q]);
                    lo[q] = f2bf(e[q] - bf2f(hi[q]));
                }
                *reinterpret_cast<uint2*>(&As[row * PLDA + v * 4]) =
                    make_uint2(hi[0] | ((uint32_t)hi[1] << 16), hi[2] | ((uint32_t)hi[3] << 16));
                *reinterpret_cast<uint2*>(&Al[row * PLDA + v * 4]) =
                    make_uint2(lo[0] | ((uint32_t)lo[1] << 16), lo[2] | ((uint32_t)lo[3] << 16));
            }
        }
        __syncthreads();
        for (int k0 = 0; k0 < kw; k0 += 32) {
            bf8v b[CT], a[4];
#pragma unroll
            for (int j = 0; j < CT; j++)
                b[j] = __builtin_bit_cast(bf8v, *reinterpret_cast<const uint4*>(Bt + (long)(nw + j * 16 + r16) * K + kc + k0 + kq));
#pragma unroll
            for (int i = 0; i < 4; i++)
                a[i] = __builtin_bit_cast(bf8v, *reinterpret_cast<const uint4*>(&As[(i * 16 + r16) * PLDA + k0 + kq]));
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < CT; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
            if (SPLIT) {
#pragma unroll
                for (int i = 0; i < 4; i++)
                    a[i] = __builtin_bit_cast(bf8v, *reinterpret_cast<const uint4*>(&Al[(i * 16 + r16) * PLDA + k0 + kq]));
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < CT; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        }
    }

    float bcol[CT];
#pragma unroll
    for (int j = 0; j < CT; j++) bcol[j] = bias[nw + j * 16 + r16];

    if (EPI == EPI_STORE || EPI == EPI_GELU) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + i * 16 + (lane >> 4) * 4 + r;
                if (m >= M) continue;
#pragma unroll
                for (int j = 0; j < CT; j++) {
                    float v = acc[i][j][r] + bcol[j];
                    if (EPI == EPI_GELU) v = gelu(v);
                    const long o = (long)m * N + nw + j * 16 + r16;
                    if (SPLIT) static_cast<float*>(Cv)[o] = v;
                    else static_cast<uint16_t*>(Cv)[o] = f2bf(v);
                }
            }
        return;
    }
    // residual: x = X + acc + bias, kept in the accumulators
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = m0 + i * 16 + (lane >> 4) * 4 + r;
            if (m >= M) continue;
#pragma unroll
            for (int j = 0; j < CT; j++) {
                float* p = X + (long)m * N + nw + j * 16 + r16;
                const float v = *p + acc[i][j][r] + bcol[j];
                *p = v;
                acc[i][j][r] = v;
            }
        }
    if (EPI != EPI_RESID_LN) return;

    // LayerNorm of the updated rows (cpu.rs:49-65): mean, then biased
    // variance around it, eps 1e-5 inside the sqrt. A row's N values are
    // spread over 16 lanes x CT per wave and over the 4 waves.
    float mean[4][4], inv[4][4];
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                float s = 0.f;
#pragma unroll
                for (int j = 0; j < CT; j++) {
                    const float v = acc[i][j][r];
                    s += pass == 0 ? v : (v - mean[i][r]) * (v - mean[i][r]);
                }
                s += __shfl_xor(s, 1, 64);
                s += __shfl_xor(s, 2, 64);
                s += __shfl_xor(s, 4, 64);
                s += __shfl_xor(s, 8, 64);
                if (r16 == 0) red[wave][i * 16 + (lane >> 4) * 4 + r] = s;
            }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = i * 16 + (lane >> 4) * 4 + r;
                const float t = red[0][row] + red[1][row] + red[2][row] + red[3][row];
                if (pass == 0) mean[i][r] = t / (float)N;
                else inv[i][r] = 1.0f / sqrtf(t / (float)N + 1e-5f);
            }
        __syncthreads();
    }
    float gc[CT], bc[CT];
#pragma unroll
    for (int j = 0; j < CT; j++) {
        gc[j] = lng[nw + j * 16 + r16];
        bc[j] = lnb[nw + j * 16 + r16];
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = m0 + i * 16 + (lane >> 4) * 4 + r;
            if (m >= M) continue;
#pragma unroll
            for (int j = 0; j < CT; j++) {
                const float h = (acc[i][j][r] - mean[i][r]) * inv[i][r] * gc[j] + bc[j];
                const long o = (long)m * N + nw + j * 16 + r16;
                if (SPLIT) static_cast<float*>(Hv)[o] = h;
                else static_cast<uint16_t*>(Hv)[o] = f2bf(h);
            }
        }
}

template <int CT, bool SPLIT>
static void launch_ct(hipStream_t s, dim3 grid, int epi, int M, int N, int K, const void* A, const uint16_t* Bt,
                      const float* bias, void* C, float* X, const float* g, const float* b, void* H) {
    switch (epi) {
        case EPI_STORE: hipLaunchKernelGGL((k_panel_bf16<CT, EPI_STORE, SPLIT>), grid, dim3(256), 0, s, M, N, K, A, Bt, bias, C, X, g, b, H); break;
        case EPI_GELU: hipLaunchKernelGGL((k_panel_bf16<CT, EPI_GELU, SPLIT>), grid, dim3(256), 0, s, M, N, K, A, Bt, bias, C, X, g, b, H); break;
        case EPI_RESID: hipLaunchKernelGGL((k_panel_bf16<CT, EPI_RESID, SPLIT>), grid, dim3(256), 0, s, M, N, K, A, Bt, bias, C, X, g, b, H); break;
        default: hipLaunchKernelGGL((k_panel_bf16<CT, EPI_RESID_LN, SPLIT>), grid, dim3(256), 0, s, M, N, K, A, Bt, bias, C, X, g, b, H); break;
    }
}

bool gemm_bf16_panel_ok(int N, int K) { return N % 64 == 0 && K % 32 == 0 && (N / 64 <= 8 || (N / 64) % 4 == 0); }

template <bool SPLIT>
static void panel(hipStream_t s, int M, int N, int K, const void* A, const uint16_t* Bt, const float* bias, void* C,
                  float* X, int epi, const float* ln_g, const float* ln_b, void* H) {
    if (M <= 0) return;
    if (!gemm_bf16_panel_ok(N, K)) fail(J2K_ERR_LIMIT, "gemm_bf16_panel: N must be a multiple of 64 and K of 32");
    if (((uintptr_t)A & 15) || ((uintptr_t)Bt & 15)) fail(J2K_ERR_INVALID, "gemm_bf16_panel: operands must be 16-byte aligned");
    int ct = N / 64, gy = 1;
    if (ct > 8) {
        gy = ct / 4;
        ct = 4;
    }
    if (epi == EPI_RESID_LN && gy != 1) fail(J2K_ERR_LIMIT, "gemm_bf16_panel: the LayerNorm epilogue needs N <= 512");
    dim3 grid(cdiv(M, PBM), gy);
    switch (ct) {
        case 1: launch_ct<1, SPLIT>(s, grid, epi, M, N, K, A, Bt, bias, C, X, ln_g, ln_b, H); break;
        case 2: launch_ct<2, SPLIT>(s, grid, epi, M, N, K, A, Bt, bias, C, X, ln_g, ln_b, H); break;
        case 3: launch_ct<3, SPLIT>(s, grid, epi, M, N, K, A, Bt, bias, C, X, ln_g, ln_b, H); break;
        case 4: launch_ct<4, SPLIT>(s, grid, epi, M, N, K, A, Bt, bias, C, X, ln_g, ln_b, H); break;
        case 5: launch_ct<5, SPLIT>(s, grid, epi, M, N, K, A, Bt, bias, C, X, ln_g, ln_b, H); break;
        case 6: launch_ct<6, SPLIT>(s, grid, epi, M, N, K, A, Bt, bias, C, X, ln_g, ln_b, H); break;
        case 7: launch_ct<7, SPLIT>(s, grid, epi, M, N, K, A, Bt, bias, C, X, ln_g, ln_b, H); break;
        default: launch_ct<8, SPLIT>(s, grid, epi, M, N, K, A, Bt, bias, C, X, ln_g, ln_b, H); break;
    }
    LAUNCHCHK("gemm_bf16_panel");
}

void gemm_bf16_panel(hipStream_t s, int M, int N, int K, const uint16_t* A, const uint16_t* Bt, const float* bias,
                     uint16_t* C, float* X, int epi, const float* ln_g, const float* ln_b, uint16_t* H) {
    panel<false>(s, M, N, K, A, Bt, bias, C, X, epi, ln_g, ln_b, H);
}

void gemm_bf16x2_panel(hipStream_t s, int M, int N, int K, const float* A, const uint16_t* Bt, const float* bias,
                       float* C, float* X, int epi, const float* ln_g, const float* ln_b, float* H) {
    panel<true>(s, M, N, K, A, Bt, bias, C, X, epi, ln_g, ln_b, H);
}

}  // namespace j2k
