// f32 GEMMs on the gfx950 matrix cores (v_mfma_f32_16x16x4_f32).
//
// Block tile 64x64, K step 16, 256 threads = 4 waves in a 2x2 grid, each wave
// a 32x32 sub-tile = 2x2 MFMA 16x16 tiles. A is staged k-major in LDS so
// each MFMA operand fetch is 16 consecutive floats per k row. Operand
// layout of the 16x16x4 instruction: lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; result element r of lane l is D[4*(l/16)+r][l%16].
//
// Loaders and epilogues are functors so one kernel body serves the plain
// GEMMs (activations x weights), the patch embed and the implicit-GEMM conv.
//
// A launch of fewer than 128 tiles (the reference forward's 256-token
// products and its CellHead convolutions) splits K over blockIdx.z when the
// caller gives a workspace: each split stores its partial tile, and a second
// kernel sums the partials in split order and applies the epilogue, so the
// result does not depend on scheduling. Larger launches are one pass.
#include <algorithm>

#include "common.h"
#include "kernels.h"

namespace j2k {

typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16, PAD = 4;

// Row-major A [M, K] with leading dimension lda.
struct LoadA {
    const float* A;
    int lda, M, K;
    bool vec;
    __device__ float4 operator()(int m, int k) const {
        float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
        if (m >= M) return r;
        const float* p = A + (long)m * lda + k;
        if (vec && k + 3 < K) return *reinterpret_cast<const float4*>(p);
        if (k < K) r.x = p[0];
        if (k + 1 < K) r.y = p[1];
        if (k + 2 < K) r.z = p[2];
        if (k + 3 < K) r.w = p[3];
        return r;
    }
};

// Row-major B [K, N]: 4 consecutive n of row k.
struct LoadB {
    const float* B;
    int ldb, K, N;
    bool vec;
    __device__ float4 operator()(int k, int n) const {
        float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
        if (k >= K) return r;
        const float* p = B + (long)k * ldb + n;
        if (vec && n + 3 < N) return *reinterpret_cast<const float4*>(p);
        if (n < N) r.x = p[0];
        if (n + 1 < N) r.y = p[1];
        if (n + 2 < N) r.z = p[2];
        if (n + 3 < N) r.w = p[3];
        return r;
    }
};

// im2col view of a [B, cin, h, w] input for the conv GEMM C[co, pixel]:
// B-operand element (k = ci*ks*ks + ky*ks + kx, n = b*h*w + y*w + x).
struct LoadIm2col {
    const float* in;
    const float* in2;
    int cin, cin1, h, w, ks, K, N;
    long bstride, bstride2;
    __device__ float at(int k, int n) const {
        if (k >= K || n >= N) return 0.f;
        int hw = h * w, b = n / hw, p = n - b * hw, y = p / w, x = p - y * w;
        int kk = ks * ks, ci = k / kk, r = k - ci * kk, ky = r / ks, kx = r - ky * ks;
        int iy = y + ky - ks / 2, ix = x + kx - ks / 2;
        if (iy < 0 || ix < 0 || iy >= h || ix >= w) return 0.f;
        if (ci < cin1) return in[b * bstride + (long)ci * hw + iy * w + ix];
        return in2[b * bstride2 + (long)(ci - cin1) * hw + iy * w + ix];
    }
    __device__ float4 operator()(int k, int n) const {
        return make_float4(at(k, n), at(k, n + 1), at(k, n + 2), at(k, n + 3));
    }
};

struct EpiPlain {
    const float* bias;
    float* C;
    int ldc, mode;
    __device__ void operator()(int m, int n, float v) const {
        v += bias ? bias[n] : 0.f;
        float* p = C + (long)m * ldc + n;
        if (mode == EPI_GELU) v = gelu(v);
        if (mode == EPI_RESID) v += *p;
        *p = v;
    }
};

// cpu.rs:271-280: x = patches @ We + be, then + pos[t], then + mvec.
struct EpiEmbed {
    const float* bias;
    const float* pos;
    const float* mvec;
    float* C;
    int N, tokens;
    __device__ void operator()(int m, int n, float v) const {
        v += bias[n];
        v += pos[(long)(m % tokens) * N + n];
        v += mvec[n];
        C[(long)m * N + n] = v;
    }
};

// Conv output C[co, pixel] -> out[b, co, y, x]; optional GELU; optional
// out = add_src[b, co, y, x] + value.
struct EpiConv {
    const float* bias;
    float* out;
    const float* add_src;
    int cout, hw, gelu_on;
    __device__ void operator()(int co, int n, float v) const {
        v += bias[co];
        if (gelu_on) v = gelu(v);
        int b = n / hw, p = n - b * hw;
        long o = ((long)b * cout + co) * hw + p;
        if (add_src) v = add_src[o] + v;
        out[o] = v;
    }
};

// A is either a plain row-major matrix (LoadA) or, for the conv, the weight
// [cout, K]; B is a plain weight (LoadB) or the im2col view. With part,
// split z covers k in [z kc, (z + 1) kc) and stores its raw tile to
// part[z][M][N] instead of applying the epilogue.
template <class AL, class BL, class EP>
__global__ __launch_bounds__(256) void k_gemm_f32(int M, int N, int K, AL la, BL lb, EP ep, int kc,
                                                  float* __restrict__ part) {
    __shared__ __align__(16) float As[BK][BM + PAD];
    __shared__ __align__(16) float Bs[BK][BN + PAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f4v acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};

    const int ar = tid >> 2, ak = (tid & 3) * 4;    // A tile: row, 4 consecutive k
    const int bk = tid >> 4, bn = (tid & 15) * 4;   // B tile: k row, 4 consecutive n
    const int kb = blockIdx.z * kc, ke = min(K, kb + kc);  // kc is a multiple of BK
    for (int k0 = kb; k0 < ke; k0 += BK) {
        float4 a = la(m0 + ar, k0 + ak);
        float4 b = lb(k0 + bk, n0 + bn);
        As[ak + 0][ar] = a.x;
        As[ak + 1][ar] = a.y;
        As[ak + 2][ar] = a.z;
        As[ak + 3][ar] = a.w;
        *reinterpret_cast<float4*>(&Bs[bk][bn]) = b;
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4), c = lane & 15;
            float av[2], bv[2];
#pragma unroll
            for (int i = 0; i < 2; i++) av[i] = As[kr][wm + i * 16 + c];
#pragma unroll
            for (int j = 0; j < 2; j++) bv[j] = Bs[kr][wn + j * 16 + c];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int n = n0 + wn + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
                if (m >= M || n >= N) continue;
                if (part) part[((long)blockIdx.z * M + m) * N + n] = acc[i][j][r];
                else ep(m, n, acc[i][j][r]);
            }
        }
}

// The split-K partials of one launch summed in split order, then the epilogue.
template <class EP>
__global__ __launch_bounds__(256) void k_splitk_reduce(int M, int N, int splits, const float* __restrict__ part,
                                                       EP ep) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x, MN = (long)M * N;
    if (i >= MN) return;
    float v = part[i];
    for (int z = 1; z < splits; z++) v += part[z * MN + i];
    ep((int)(i / N), (int)(i % N), v);
}

// Splits of K for a launch of `tiles` tiles (1: one pass), each a multiple
// of BK and at least 32 deep, so that about 256 blocks run; *kc receives the
// depth of one split. Needs splits * M * N floats of workspace.
static int splitk_of(int tiles, int K, int* kc) {
    int s = 1;
    if (tiles < 128 && K >= 64) s = std::max(1, std::min(256 / tiles, K / 32));
    *kc = (cdiv(K, s) + BK - 1) / BK * BK;
    return cdiv(K, *kc);
}

template <class AL, class BL, class EP>
static void launch_gemm(hipStream_t s, int M, int N, int K, const AL& la, const BL& lb, const EP& ep, float* ws,
                        size_t ws_floats, const char* what) {
    const int tm = cdiv(M, BM), tn = cdiv(N, BN);
    int kc = K;
    int splits = splitk_of(tm * tn, K, &kc);
    if (!ws || (size_t)splits * M * N > ws_floats) splits = 1, kc = (K + BK - 1) / BK * BK;
    dim3 grid(tn, tm, splits);
    hipLaunchKernelGGL((k_gemm_f32<AL, BL, EP>), grid, dim3(256), 0, s, M, N, K, la, lb, ep, kc,
                       splits > 1 ? ws : nullptr);
    if (splits > 1) {
        const long MN = (long)M * N;
        hipLaunchKernelGGL((k_splitk_reduce<EP>), dim3(cdiv(MN, 256)), dim3(256), 0, s, M, N, splits, ws, ep);
    }
    LAUNCHCHK(what);
}

size_t gemm_f32_workspace(int M, int N, int K) {
    int kc = K;
    const int splits = splitk_of(cdiv(M, BM) * cdiv(N, BN), K, &kc);
    return splits > 1 ? (size_t)splits * M * N : 0;
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

void gemm_f32(hipStream_t s, int M, int N, int K, const float* A, int lda, const float* B, const float* bias, float* C,
              int ldc, int epi, float* ws, size_t ws_floats) {
    if (M <= 0 || N <= 0) return;
    LoadA la{A, lda, M, K, aligned16(A) && lda % 4 == 0};
    LoadB lb{B, N, K, N, aligned16(B) && N % 4 == 0};
    EpiPlain ep{bias, C, ldc, epi};
    launch_gemm(s, M, N, K, la, lb, ep, ws, ws_floats, "gemm_f32");
}

void gemm_embed_p(hipStream_t s, int M, int N, int K, const float* A, const float* B, const float* bias,
                  const float* pos, int tokens, const float* mvec, float* C, float* ws, size_t ws_floats) {
    LoadA la{A, K, M, K, aligned16(A) && K % 4 == 0};
    LoadB lb{B, N, K, N, aligned16(B) && N % 4 == 0};
    EpiEmbed ep{bias, pos, mvec, C, N, tokens};
    launch_gemm(s, M, N, K, la, lb, ep, ws, ws_floats, "gemm_embed_p");
}

void conv_f32(hipStream_t s, int B, int cin, int h, int w, const float* in, const float* in2, int cin1,
              const float* wt, const float* bias, int cout, int k, float* out, int gelu_on, const float* add_src,
              float* ws, size_t ws_floats) {
    const int K = cin * k * k, N = B * h * w, M = cout;
    if (!in2) cin1 = cin;
    LoadA la{wt, K, M, K, aligned16(wt) && K % 4 == 0};
    LoadIm2col lb{in, in2, cin, cin1, h, w, k, K, N, (long)cin1 * h * w, (long)(cin - cin1) * h * w};
    EpiConv ep{bias, out, add_src, cout, h * w, gelu_on};
    launch_gemm(s, M, N, K, la, lb, ep, ws, ws_floats, "conv_f32");
}

}  // namespace j2k
