// bf16 GEMMs on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16), f32
// accumulate, fused bias / GELU / residual epilogue.
//
// Operands: A [M, K] bf16 row-major (activations), Bt [N, K] bf16 (the
// weight stored transposed at load), so both are K-contiguous. For the
// 16x16x32 instruction lane l supplies A[l%16][8*(l/16) .. +8) and
// Bt[l%16][8*(l/16) .. +8) as one 16-byte vector each; result element r of
// lane l is D[4*(l/16)+r][l%16].
//
// Block 256 threads = 4 waves, block tile 64 x 64, each wave a 32 x 32
// sub-tile (2 x 2 MFMA tiles). Operands are read straight from global
// memory in 16-byte vectors (rows shared by two waves hit L1).
#include "common.h"
#include "kernels.h"

namespace j2k {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf8v load_bf8(const uint16_t* p, bool ok) {
    uint4 v = ok ? *reinterpret_cast<const uint4*>(p) : make_uint4(0, 0, 0, 0);
    return __builtin_bit_cast(bf8v, v);
}

__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A,
                                                   const uint16_t* __restrict__ Bt, const float* __restrict__ bias,
                                                   uint16_t* __restrict__ C, float* __restrict__ Cf, int ldc, int epi) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * 64 + (wave >> 1) * 32, n0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int r16 = lane & 15, kg = (lane >> 4) * 8;
    f4v acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};
    const uint16_t* ap[2];
    const uint16_t* bp[2];
    bool aok[2], bok[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int m = m0 + i * 16 + r16, n = n0 + i * 16 + r16;
        aok[i] = m < M;
        bok[i] = n < N;
        ap[i] = A + (long)(aok[i] ? m : 0) * K + kg;
        bp[i] = Bt + (long)(bok[i] ? n : 0) * K + kg;
    }
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf8v a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            a[i] = load_bf8(ap[i] + k0, aok[i]);
            b[i] = load_bf8(bp[i] + k0, bok[i]);
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int n = n0 + j * 16 + r16;
            if (n >= N) continue;
            const float bn = bias ? bias[n] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + i * 16 + (lane >> 4) * 4 + r;
                if (m >= M) continue;
                float v = acc[i][j][r] + bn;
                if (epi == EPI_RESID) {
                    Cf[(long)m * ldc + n] += v;
                } else {
                    if (epi == EPI_GELU) v = gelu(v);
                    C[(long)m * ldc + n] = f2bf(v);
                }
            }
        }
}

void gemm_bf16(hipStream_t s, int M, int N, int K, const uint16_t* A, const uint16_t* Bt, const float* bias, uint16_t* C,
               float* Cf, int ldc, int epi) {
    if (M <= 0 || N <= 0) return;
    if (K % 32) fail(J2K_ERR_LIMIT, "gemm_bf16: K must be a multiple of 32");
    if (((uintptr_t)A & 15) || ((uintptr_t)Bt & 15)) fail(J2K_ERR_INVALID, "gemm_bf16: operands must be 16-byte aligned");
    dim3 grid(cdiv(N, 64), cdiv(M, 64));
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s, M, N, K, A, Bt, bias, C, Cf, ldc, epi);
    LAUNCHCHK("gemm_bf16");
}

}  // namespace j2k
