// Tiled bf16 GEMM for the trunk's products (M = tokens, N and K a few
// hundred) on v_mfma_f32_16x16x32_bf16 with f32 accumulation.
//
// Block tile BM x 128 (BM = 64 or 128, chosen per shape by the autotuner),
// 256 threads = 4 waves; each wave owns 64 rows and 128 / (4*64/BM) columns
// of 16 x 16 MFMA tiles. K advances in chunks of 64: a chunk of A and of Bt
// (both K-contiguous) sits in LDS, rows padded by 8 bf16 so the 16-row
// fragment reads of a wave hit distinct banks, while the next chunk is
// already loading into registers.
//
// Blocks map to tiles XCD by XCD (xcd_remap): each XCD takes a contiguous
// run of tiles, so the column tiles of one row panel land on one XCD and
// share its L2.
//
// SPLIT: A is f32 and is staged as a bf16 pair hi = bf16(a), lo = bf16(a-hi);
// every tile takes two MFMAs against the same bf16 weight fragment, so the
// activations keep about 16 mantissa bits and the product carries only the
// weights' bf16 rounding. Outputs are f32 then.
//
// Epilogues: store, GELU + store, residual add into the f32 stream X, and
// residual add followed by the LayerNorm of the updated rows (N == 128: the
// block owns whole rows), which writes the next GEMM's input H.
#include "common.h"
#include "kernels.h"
#include "lnstats.h"

namespace j2k {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

constexpr int TBN = 128, TBK = 64, TLD = TBK + 8;

// Epilogue of a BM x 128 tile whose accumulators acc[i][j] (row wm + i*16 +
// 4*(lane/16) + r, column wn + j*16 + lane%16) are spread over the block's
// four waves: v = acc * alpha + bias, then store / GELU + store / residual
// add into X / residual add and the LayerNorm of the updated rows. alpha is
// a * colscale[n] with colscale per output column (nullptr: 1), and with RS
// also times the row's scale rowmax[m] / 448 (1 for a row of zeros). C and H
// are f32 (OUTF32) or bf16.
__device__ __forceinline__ float row_scale(float amax) { return amax > 0.f ? amax / 448.0f : 1.0f; }

template <int EPI, bool OUTF32, int BM, int J, bool RS = false>
__device__ __forceinline__ void tile_epilogue(f4v (&acc)[4][J], float a, const float* __restrict__ colscale,
                                              const float* __restrict__ rowmax, int M,
                                              int N, int m0, int n0, int wm, int wn, int lane,
                                              const float* __restrict__ bias, void* __restrict__ Cv,
                                              float* __restrict__ X, const float* __restrict__ lng,
                                              const float* __restrict__ lnb, void* __restrict__ Hv,
                                              float (*red)[BM], float* mrow, float* irow) {
    const int r16 = lane & 15;
    float bcol[J], acol[J];
    bool colok[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        const int n = n0 + wn + j * 16 + r16;
        colok[j] = n < N;
        bcol[j] = colok[j] ? bias[n] : 0.f;
        acol[j] = colscale && colok[j] ? a * colscale[n] : a;
    }
    if (EPI == EPI_STORE || EPI == EPI_GELU) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
                if (m >= M) continue;
                const float rs = RS ? row_scale(rowmax[m]) : 1.0f;
#pragma unroll
                for (int j = 0; j < J; j++) {
                    if (!colok[j]) continue;
                    float v = acc[i][j][r] * (RS ? rs * acol[j] : acol[j]) + bcol[j];
                    if (EPI == EPI_GELU) v = gelu_fast(v);
                    const long o = (long)m * N + n0 + wn + j * 16 + r16;
                    if (OUTF32) static_cast<float*>(Cv)[o] = v;
                    else static_cast<uint16_t*>(Cv)[o] = f2bf(v);
                }
            }
        return;
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
            if (m >= M) continue;
            const float rs = RS ? row_scale(rowmax[m]) : 1.0f;
#pragma unroll
            for (int j = 0; j < J; j++) {
                if (!colok[j]) continue;
                float* p = X + (long)m * N + n0 + wn + j * 16 + r16;
                const float v = *p + acc[i][j][r] * (RS ? rs * acol[j] : acol[j]) + bcol[j];
                *p = v;
                acc[i][j][r] = v;
            }
        }
    if (EPI != EPI_RESID_LN) return;

    // LayerNorm of the updated rows (cpu.rs:49-65); N == 128, so the block
    // owns whole rows, 8 column tiles each.
    ln_row_stats<BM, J>(acc, wm, wn, lane, red, mrow, irow);
    float gc[J], bc[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        gc[j] = lng[wn + j * 16 + r16];
        bc[j] = lnb[wn + j * 16 + r16];
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = wm + i * 16 + (lane >> 4) * 4 + r, m = m0 + row;
            if (m >= M) continue;
            const float mu = mrow[row], iv = irow[row];
#pragma unroll
            for (int j = 0; j < J; j++) {
                const float h = (acc[i][j][r] - mu) * iv * gc[j] + bc[j];
                const long o = (long)m * N + wn + j * 16 + r16;
                if (OUTF32) static_cast<float*>(Hv)[o] = h;
                else static_cast<uint16_t*>(Hv)[o] = f2bf(h);
            }
        }
}

template <int EPI, bool SPLIT, int BM>
__global__ __launch_bounds__(256, 2) void k_tile_bf16(int M, int N, int K, const void* __restrict__ Av,
                                                      const uint16_t* __restrict__ Bt, const float* __restrict__ bias,
                                                      void* __restrict__ Cv, float* __restrict__ X,
                                                      const float* __restrict__ lng, const float* __restrict__ lnb,
                                                      void* __restrict__ Hv, int nt) {
    constexpr int WMW = BM / 64, WNW = 4 / WMW;  // waves along rows / columns
    constexpr int WN = TBN / WNW, J = WN / 16;   // columns per wave, column tiles per wave
    __shared__ __align__(16) uint16_t Ah[BM * TLD];
    __shared__ __align__(16) uint16_t Al[SPLIT ? BM * TLD : 8];
    __shared__ __align__(16) uint16_t Bs[TBN * TLD];
    __shared__ float red[TBN / 16][BM], mrow[BM], irow[BM];

    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int m0 = (t / nt) * BM, n0 = (t % nt) * TBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave / WNW) * 64, wn = (wave % WNW) * WN;
    const int r16 = lane & 15, kq = (lane >> 4) * 8;

    f4v acc[4][J];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < J; j++) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};

    constexpr int NA = (SPLIT ? 16 : 8) * BM / 256;  // 16-byte vectors of A per thread per chunk
    uint4 pa[NA], pb[4];
    auto load_chunk = [&](int kc) {
        if (!SPLIT) {
            const uint16_t* A = static_cast<const uint16_t*>(Av);
#pragma unroll
            for (int q = 0; q < NA; q++) {
                const int v = tid + 256 * q, row = v >> 3, k = kc + (v & 7) * 8, m = m0 + row;
                pa[q] = (m < M && k < K) ? *reinterpret_cast<const uint4*>(A + (long)m * K + k) : make_uint4(0, 0, 0, 0);
            }
        } else {
            const float* A = static_cast<const float*>(Av);
#pragma unroll
            for (int q = 0; q < NA; q++) {
                const int v = tid + 256 * q, row = v >> 4, k = kc + (v & 15) * 4, m = m0 + row;
                pa[q] = (m < M && k < K) ? *reinterpret_cast<const uint4*>(A + (long)m * K + k) : make_uint4(0, 0, 0, 0);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int v = tid + 256 * q, row = v >> 3, k = kc + (v & 7) * 8, n = n0 + row;
            pb[q] = (n < N && k < K) ? *reinterpret_cast<const uint4*>(Bt + (long)n * K + k) : make_uint4(0, 0, 0, 0);
        }
    };
    auto store_chunk = [&]() {
        if (!SPLIT) {
#pragma unroll
            for (int q = 0; q < NA; q++) {
                const int v = tid + 256 * q, row = v >> 3, kv = v & 7;
                *reinterpret_cast<uint4*>(&Ah[row * TLD + kv * 8]) = pa[q];
            }
        } else {
#pragma unroll
            for (int q = 0; q < NA; q++) {
                const int v = tid + 256 * q, row = v >> 4, kv = v & 15;
                const float f0 = __uint_as_float(pa[q].x), f1 = __uint_as_float(pa[q].y);
                const float f2 = __uint_as_float(pa[q].z), f3 = __uint_as_float(pa[q].w);
                const uint32_t h01 = f2bf2(f0, f1), h23 = f2bf2(f2, f3);
                const uint32_t l01 = f2bf2(f0 - __uint_as_float(h01 << 16), f1 - __uint_as_float(h01 & 0xffff0000u));
                const uint32_t l23 = f2bf2(f2 - __uint_as_float(h23 << 16), f3 - __uint_as_float(h23 & 0xffff0000u));
                *reinterpret_cast<uint2*>(&Ah[row * TLD + kv * 4]) = make_uint2(h01, h23);
                *reinterpret_cast<uint2*>(&Al[row * TLD + kv * 4]) = make_uint2(l01, l23);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int v = tid + 256 * q, row = v >> 3, kv = v & 7;
            *reinterpret_cast<uint4*>(&Bs[row * TLD + kv * 8]) = pb[q];
        }
    };

    load_chunk(0);
    for (int kc = 0; kc < K; kc += TBK) {
        __syncthreads();  // the previous chunk's fragments have been read
        store_chunk();
        __syncthreads();
        if (kc + TBK < K) load_chunk(kc + TBK);
#pragma unroll
        for (int ks = 0; ks < TBK; ks += 32) {
            bf8v a[4], bb[J];
#pragma unroll
            for (int j = 0; j < J; j++)
                bb[j] = __builtin_bit_cast(bf8v, *reinterpret_cast<const uint4*>(&Bs[(wn + j * 16 + r16) * TLD + ks + kq]));
#pragma unroll
            for (int i = 0; i < 4; i++)
                a[i] = __builtin_bit_cast(bf8v, *reinterpret_cast<const uint4*>(&Ah[(wm + i * 16 + r16) * TLD + ks + kq]));
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < J; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bb[j], acc[i][j], 0, 0, 0);
            if (SPLIT) {
#pragma unroll
                for (int i = 0; i < 4; i++)
                    a[i] = __builtin_bit_cast(bf8v, *reinterpret_cast<const uint4*>(&Al[(wm + i * 16 + r16) * TLD + ks + kq]));
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < J; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bb[j], acc[i][j], 0, 0, 0);
            }
        }
    }

    tile_epilogue<EPI, SPLIT, BM, J>(acc, 1.0f, nullptr, nullptr, M, N, m0, n0, wm, wn, lane, bias, Cv, X, lng, lnb, Hv, red,
                                     mrow, irow);
}

// FP8 form: A is f32 and is quantized while staged, q = clamp(a * inv_m,
// +-448) rounded to OCP E4M3 (v_cvt_pk_fp8_f32), with each row's (token's)
// scale sa_m = rowmax[m] / 448 from its measured maximum |A[m, :]| (written by
// k_row_amax before this launch; 1 for a row of zeros, inv_m = 1 / sa_m
// computed as 448 / rowmax[m]): a token's quantization never depends on the
// other tokens of a pack. Bt holds E4M3 weight codes [N, K];
// v_mfma_f32_16x16x32_fp8_fp8 accumulates the exact code products in f32,
// and the epilogue rescales by sa_m * colscale[n] (the weights' per-tensor
// scale, per output column because the fused qkv matrix joins three
// tensors). Same tiling as the bf16 kernel with 8-byte fragments.
constexpr int T8LD = TBK + 16;  // bytes per LDS row

template <int EPI, int BM>
__global__ __launch_bounds__(256, 2) void k_tile_f8(int M, int N, int K, const float* __restrict__ A,
                                                    const float* __restrict__ rowmax, const uint8_t* __restrict__ Bt,
                                                    const float* __restrict__ colscale, const float* __restrict__ bias,
                                                    float* __restrict__ C, float* __restrict__ X,
                                                    const float* __restrict__ lng, const float* __restrict__ lnb,
                                                    float* __restrict__ H, int nt) {
    constexpr int WMW = BM / 64, WNW = 4 / WMW;
    constexpr int WN = TBN / WNW, J = WN / 16;
    __shared__ __align__(16) uint8_t As[BM * T8LD];
    __shared__ __align__(16) uint8_t Bs[TBN * T8LD];
    __shared__ float red[TBN / 16][BM], mrow[BM], irow[BM];

    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int m0 = (t / nt) * BM, n0 = (t % nt) * TBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave / WNW) * 64, wn = (wave % WNW) * WN;
    const int r16 = lane & 15, kq = (lane >> 4) * 8;

    f4v acc[4][J];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < J; j++) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};

    constexpr int NA = 16 * BM / 256;  // float4 of A per thread per chunk
    float4 pa[NA];
    uint4 pb[2];
    float inv_sa[NA];  // 1 / scale of the rows this thread stages (the same rows in every chunk)
#pragma unroll
    for (int q = 0; q < NA; q++) {
        const int m = m0 + ((tid + 256 * q) >> 4);
        const float am = m < M ? rowmax[m] : 0.f;
        inv_sa[q] = am > 0.f ? 448.0f / am : 1.0f;
    }
#define J2K_F8_LOAD(kc)                                                                                        \
    {                                                                                                          \
        _Pragma("unroll") for (int q = 0; q < NA; q++) {                                                       \
            const int v = tid + 256 * q, row = v >> 4, k = (kc) + (v & 15) * 4, m = m0 + row;                  \
            pa[q] = (m < M && k < K) ? *reinterpret_cast<const float4*>(A + (long)m * K + k)                    \
                                     : make_float4(0.f, 0.f, 0.f, 0.f);                                        \
        }                                                                                                      \
        _Pragma("unroll") for (int q = 0; q < 2; q++) {                                                        \
            const int v = tid + 256 * q, row = v >> 2, k = (kc) + (v & 3) * 16, n = n0 + row;                  \
            pb[q] = (n < N && k < K) ? *reinterpret_cast<const uint4*>(Bt + (long)n * K + k) : make_uint4(0, 0, 0, 0); \
        }                                                                                                      \
    }
    J2K_F8_LOAD(0)
    for (int kc = 0; kc < K; kc += TBK) {
        __syncthreads();  // the previous chunk's fragments have been read
#pragma unroll
        for (int q = 0; q < NA; q++) {
            const int v = tid + 256 * q, row = v >> 4, kv = v & 15;
            const float e0 = fminf(fmaxf(pa[q].x * inv_sa[q], -448.f), 448.f);
            const float e1 = fminf(fmaxf(pa[q].y * inv_sa[q], -448.f), 448.f);
            const float e2 = fminf(fmaxf(pa[q].z * inv_sa[q], -448.f), 448.f);
            const float e3 = fminf(fmaxf(pa[q].w * inv_sa[q], -448.f), 448.f);
            int w = __builtin_amdgcn_cvt_pk_fp8_f32(e0, e1, 0, false);
            w = __builtin_amdgcn_cvt_pk_fp8_f32(e2, e3, w, true);
            *reinterpret_cast<int*>(&As[row * T8LD + kv * 4]) = w;
        }
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int v = tid + 256 * q, row = v >> 2, kv = v & 3;
            *reinterpret_cast<uint4*>(&Bs[row * T8LD + kv * 16]) = pb[q];
        }
        __syncthreads();
        if (kc + TBK < K) J2K_F8_LOAD(kc + TBK)
#pragma unroll
        for (int ks = 0; ks < TBK; ks += 32) {
            long a8[4], b8[J];
#pragma unroll
            for (int j = 0; j < J; j++) b8[j] = *reinterpret_cast<const long*>(&Bs[(wn + j * 16 + r16) * T8LD + ks + kq]);
#pragma unroll
            for (int i = 0; i < 4; i++) a8[i] = *reinterpret_cast<const long*>(&As[(wm + i * 16 + r16) * T8LD + ks + kq]);
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < J; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a8[i], b8[j], acc[i][j], 0, 0, 0);
        }
    }
#undef J2K_F8_LOAD
    tile_epilogue<EPI, true, BM, J, true>(acc, 1.0f, colscale, rowmax, M, N, m0, n0, wm, wn, lane, bias, C, X, lng, lnb,
                                          H, red, mrow, irow);
}

bool gemm_tile_ok(int N, int K, bool split) { return N % 16 == 0 && K % (split ? 4 : 8) == 0; }

int gemm_tile_default_bm(int M, int N) {
    // enough blocks to give every CU two at BM = 128, else halve the rows
    const long blocks128 = (long)cdiv(M, 128) * cdiv(N, TBN);
    return blocks128 >= 512 ? 128 : 64;
}

template <bool SPLIT, int BM>
static void tile_bm(hipStream_t s, int M, int N, int K, const void* A, const uint16_t* Bt, const float* bias, void* C,
                    float* X, int epi, const float* g, const float* b, void* H) {
    const int mt = cdiv(M, BM), nt = cdiv(N, TBN);
    dim3 grid(mt * nt), block(256);
    switch (epi) {
        case EPI_STORE: hipLaunchKernelGGL((k_tile_bf16<EPI_STORE, SPLIT, BM>), grid, block, 0, s, M, N, K, A, Bt, bias, C, X, g, b, H, nt); break;
        case EPI_GELU: hipLaunchKernelGGL((k_tile_bf16<EPI_GELU, SPLIT, BM>), grid, block, 0, s, M, N, K, A, Bt, bias, C, X, g, b, H, nt); break;
        case EPI_RESID: hipLaunchKernelGGL((k_tile_bf16<EPI_RESID, SPLIT, BM>), grid, block, 0, s, M, N, K, A, Bt, bias, C, X, g, b, H, nt); break;
        default: hipLaunchKernelGGL((k_tile_bf16<EPI_RESID_LN, SPLIT, BM>), grid, block, 0, s, M, N, K, A, Bt, bias, C, X, g, b, H, nt); break;
    }
}

template <bool SPLIT>
static void tile(hipStream_t s, int M, int N, int K, const void* A, const uint16_t* Bt, const float* bias, void* C,
                 float* X, int epi, const float* g, const float* b, void* H, int bm) {
    if (M <= 0) return;
    if (!gemm_tile_ok(N, K, SPLIT)) fail(J2K_ERR_LIMIT, "gemm tile: N must be a multiple of 16 and K of 8");
    if (((uintptr_t)A & 15) || ((uintptr_t)Bt & 15)) fail(J2K_ERR_INVALID, "gemm tile: operands must be 16-byte aligned");
    if (epi == EPI_RESID_LN && N != TBN) fail(J2K_ERR_LIMIT, "gemm tile: the LayerNorm epilogue needs N == 128");
    if (bm <= 0) bm = gemm_tile_default_bm(M, N);
    if (bm == 64) tile_bm<SPLIT, 64>(s, M, N, K, A, Bt, bias, C, X, epi, g, b, H);
    else tile_bm<SPLIT, 128>(s, M, N, K, A, Bt, bias, C, X, epi, g, b, H);
    LAUNCHCHK("gemm tile");
}

template <int BM>
static void tile_f8_bm(hipStream_t s, int M, int N, int K, const float* A, const float* amax_a, const uint8_t* Bt,
                       const float* cs, const float* bias, float* C, float* X, int epi, const float* g,
                       const float* b, float* H) {
    const int mt = cdiv(M, BM), nt = cdiv(N, TBN);
    dim3 grid(mt * nt), block(256);
    switch (epi) {
        case EPI_STORE: hipLaunchKernelGGL((k_tile_f8<EPI_STORE, BM>), grid, block, 0, s, M, N, K, A, amax_a, Bt, cs, bias, C, X, g, b, H, nt); break;
        case EPI_GELU: hipLaunchKernelGGL((k_tile_f8<EPI_GELU, BM>), grid, block, 0, s, M, N, K, A, amax_a, Bt, cs, bias, C, X, g, b, H, nt); break;
        case EPI_RESID: hipLaunchKernelGGL((k_tile_f8<EPI_RESID, BM>), grid, block, 0, s, M, N, K, A, amax_a, Bt, cs, bias, C, X, g, b, H, nt); break;
        default: hipLaunchKernelGGL((k_tile_f8<EPI_RESID_LN, BM>), grid, block, 0, s, M, N, K, A, amax_a, Bt, cs, bias, C, X, g, b, H, nt); break;
    }
}

// out[m] = max |A[m, :]| for the M rows of A [M, K] (K a multiple of 4, A
// 16-byte aligned): one wave per row, four rows a block.
__global__ __launch_bounds__(256) void k_row_amax(int M, int K, const float* __restrict__ A, float* __restrict__ out) {
    const int m = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (m >= M) return;
    const float4* row = reinterpret_cast<const float4*>(A + (long)m * K);
    float a = 0.f;
    for (int i = lane; i < K / 4; i += 64) {
        const float4 v = row[i];
        a = fmaxf(a, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
    }
    a = wave_max(a);
    if (lane == 0) out[m] = a;
}

void row_amax_f32(hipStream_t s, int M, int K, const float* A, float* out) {
    if (M <= 0) return;
    if ((K & 3) || ((uintptr_t)A & 15)) fail(J2K_ERR_INVALID, "row amax: rows must be 16-byte aligned");
    hipLaunchKernelGGL(k_row_amax, dim3(cdiv(M, 4)), dim3(256), 0, s, M, K, A, out);
    LAUNCHCHK("row amax");
}

void gemm_f8_tile(hipStream_t s, int M, int N, int K, const float* A, const float* amax_a, const uint8_t* Bt,
                  const float* cs, const float* bias, float* C, float* X, int epi, const float* ln_g,
                  const float* ln_b, float* H, int bm) {
    if (M <= 0) return;
    if (N % 16 || K % 16) fail(J2K_ERR_LIMIT, "gemm fp8 tile: N and K must be multiples of 16");
    if (!amax_a) fail(J2K_ERR_INVALID, "gemm fp8 tile: missing the rows' maxima");
    if (((uintptr_t)A & 15) || ((uintptr_t)Bt & 15)) fail(J2K_ERR_INVALID, "gemm fp8 tile: operands must be 16-byte aligned");
    if (epi == EPI_RESID_LN && N != TBN) fail(J2K_ERR_LIMIT, "gemm fp8 tile: the LayerNorm epilogue needs N == 128");
    if (bm <= 0) bm = gemm_tile_default_bm(M, N);
    if (bm == 64) tile_f8_bm<64>(s, M, N, K, A, amax_a, Bt, cs, bias, C, X, epi, ln_g, ln_b, H);
    else tile_f8_bm<128>(s, M, N, K, A, amax_a, Bt, cs, bias, C, X, epi, ln_g, ln_b, H);
    LAUNCHCHK("gemm fp8 tile");
}

void gemm_bf16_tile(hipStream_t s, int M, int N, int K, const uint16_t* A, const uint16_t* Bt, const float* bias,
                    uint16_t* C, float* X, int epi, const float* ln_g, const float* ln_b, uint16_t* H, int bm) {
    tile<false>(s, M, N, K, A, Bt, bias, C, X, epi, ln_g, ln_b, H, bm);
}

void gemm_bf16x2_tile(hipStream_t s, int M, int N, int K, const float* A, const uint16_t* Bt, const float* bias,
                      float* C, float* X, int epi, const float* ln_g, const float* ln_b, float* H, int bm) {
    tile<true>(s, M, N, K, A, Bt, bias, C, X, epi, ln_g, ln_b, H, bm);
}

}  // namespace j2k
