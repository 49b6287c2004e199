// Row-panel kernel for the d = 128 trunk: everything of a block
// (cpu.rs:287-313) that is local to a token, fused over 64-token panels so
// the panel's activations stay in LDS and registers:
//
//   PROJ   X += O @ Wp + bp, then H = LN2(X)         (else H comes from Hin)
//          X += ff2(GELU(ff1(H)))                    (always)
//   QKV    QKV = LN1'(X) @ Wqkv' + bqkv'             (next block's norm and qkv)
//   LNOUT  Hout = LN1'(X)                            (when the qkv GEMM runs apart)
//
// With PROJ and QKV a layer is two launches: this kernel and the attention.
// The same arithmetic, in the same order, as the tile GEMM epilogues and the
// MLP alone, so every combination gives bit-identical results and the
// autotuner may pick any of them.
//
// A block owns 64 tokens; its four waves split every 128-column product (32
// columns each, 2 x 4 tiles of v_mfma_f32_16x16x32_bf16, f32 accumulate).
// A operands (O, H, the GELU intermediate) sit in LDS as bf16 (or, SPLIT, as
// a bf16 hi + lo pair of an f32 value, two MFMAs per tile). Weights
// (K-contiguous bf16 [N, K]) stream from L2 as 16-byte fragments, loaded a
// whole 128 x 128 chunk ahead of their use. The ff-wide intermediate is made
// 128 columns at a time and consumed at once as the next K chunk of ff2.
// Row statistics of the LayerNorms in the order every served kernel uses
// (lnstats.h).
#include "common.h"
#include "kernels.h"
#include "lnstats.h"

namespace j2k {

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

constexpr int RB = 64, RD = 128, RLD = RD + 8;

__device__ __forceinline__ bf8v frag(const uint16_t* p) { return __builtin_bit_cast(bf8v, *reinterpret_cast<const uint4*>(p)); }

// Fragments of W rows n0 + wn + j*16 + r16 (j < J), K columns k0 .. k0+128.
template <int J>
__device__ __forceinline__ void load_w(bf8v (&w)[4][J], const uint16_t* Wt, long ldk, int n0, int k0, int wn, int r16,
                                       int kq) {
#pragma unroll
    for (int k4 = 0; k4 < 4; k4++)
#pragma unroll
        for (int j = 0; j < J; j++) w[k4][j] = frag(Wt + (long)(n0 + wn + j * 16 + r16) * ldk + k0 + k4 * 32 + kq);
}

// acc += A (64 x 128 in LDS, hi and lo) @ the loaded W chunk.
template <bool SPLIT, int J>
__device__ __forceinline__ void mma_panel(f4v (&acc)[4][J], const uint16_t* Ah, const uint16_t* Al,
                                          const bf8v (&w)[4][J], int r16, int kq) {
#pragma unroll
    for (int k4 = 0; k4 < 4; k4++) {
        bf8v a[4];
#pragma unroll
        for (int i = 0; i < 4; i++) a[i] = frag(&Ah[(i * 16 + r16) * RLD + k4 * 32 + kq]);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < J; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], w[k4][j], acc[i][j], 0, 0, 0);
        if (SPLIT) {
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = frag(&Al[(i * 16 + r16) * RLD + k4 * 32 + kq]);
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < J; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], w[k4][j], acc[i][j], 0, 0, 0);
        }
    }
}

template <int J>
__device__ __forceinline__ void zero(f4v (&acc)[4][J]) {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < J; j++) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};
}

// v as an A operand at (row, col): bf16, and with SPLIT the bf16 remainder.
template <bool SPLIT>
__device__ __forceinline__ void put_a(uint16_t* Ah, uint16_t* Al, int row, int col, float v) {
    const uint16_t h = f2bf(v);
    Ah[row * RLD + col] = h;
    if (SPLIT) Al[row * RLD + col] = f2bf(v - bf2f(h));
}

// 64 rows x 128 of global A (bf16, or f32 split into hi + lo) into LDS.
template <bool SPLIT, int NT>
__device__ __forceinline__ void stage_rows(const void* src, int m0, int M, uint16_t* Ah, uint16_t* Al, int tid) {
    if (!SPLIT) {
        const uint16_t* A = static_cast<const uint16_t*>(src);
#pragma unroll
        for (int q = 0; q < 1024 / NT; q++) {
            const int v = tid + NT * q, row = v >> 4, kv = v & 15, m = m0 + row;
            const uint4 val = m < M ? *reinterpret_cast<const uint4*>(A + (long)m * RD + kv * 8) : make_uint4(0, 0, 0, 0);
            *reinterpret_cast<uint4*>(&Ah[row * RLD + kv * 8]) = val;
        }
    } else {
        const float* A = static_cast<const float*>(src);
#pragma unroll
        for (int q = 0; q < 2048 / NT; q++) {
            const int v = tid + NT * q, row = v >> 5, kv = v & 31, m = m0 + row;
            const float4 f = m < M ? *reinterpret_cast<const float4*>(A + (long)m * RD + kv * 4) : make_float4(0.f, 0.f, 0.f, 0.f);
            const uint32_t h01 = f2bf2(f.x, f.y), h23 = f2bf2(f.z, f.w);
            const uint32_t l01 = f2bf2(f.x - __uint_as_float(h01 << 16), f.y - __uint_as_float(h01 & 0xffff0000u));
            const uint32_t l23 = f2bf2(f.z - __uint_as_float(h23 << 16), f.w - __uint_as_float(h23 & 0xffff0000u));
            *reinterpret_cast<uint2*>(&Ah[row * RLD + kv * 4]) = make_uint2(h01, h23);
            *reinterpret_cast<uint2*>(&Al[row * RLD + kv * 4]) = make_uint2(l01, l23);
        }
    }
}

// WAVES = 4: each wave owns 32 of a product's 128 columns (2 tiles); 8:
// 16 columns (1 tile), half the registers, twice the waves per block. The
// arithmetic of an output element is the same.
template <bool SPLIT, bool PROJ, bool QKV, bool LNOUT, int WAVES>
__global__ __launch_bounds__(64 * WAVES, WAVES / 2) void k_rows(int M, int ff, RowsArgs a) {
    constexpr int J = 8 / WAVES, NT = 64 * WAVES;
    __shared__ __align__(16) uint16_t Hh[RB * RLD];
    __shared__ __align__(16) uint16_t Hl[SPLIT ? RB * RLD : 8];
    __shared__ __align__(16) uint16_t Fh[RB * RLD];
    __shared__ __align__(16) uint16_t Fl[SPLIT ? RB * RLD : 8];
    __shared__ float red[RD / 16][RB], mrow[RB], irow[RB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * RB, wn = wave * 16 * J;
    const int r16 = lane & 15, kq = (lane >> 4) * 8, rq = (lane >> 4) * 4;
    bf8v w1[4][J], w2[4][J];
    f4v xm[4][J];  // the residual stream of the panel

    // ---- the residual stream, and H, the MLP's input
    stage_rows<SPLIT, NT>(PROJ ? a.O : a.Hin, m0, M, Hh, Hl, tid);
    if (PROJ) load_w(w1, a.Wp, RD, 0, 0, wn, r16, kq);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = m0 + i * 16 + rq + r;
#pragma unroll
            for (int j = 0; j < J; j++) xm[i][j][r] = m < M ? a.X[(long)m * RD + wn + j * 16 + r16] : 0.f;
        }
    __syncthreads();
    if (PROJ) {
        f4v acc[4][J];
        zero(acc);
        mma_panel<SPLIT, J>(acc, Hh, Hl, w1, r16, kq);
#pragma unroll
        for (int j = 0; j < J; j++) {
            const float b = a.bp[wn + j * 16 + r16];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int r = 0; r < 4; r++) xm[i][j][r] = xm[i][j][r] + acc[i][j][r] + b;
        }
        ln_row_stats<RB, J>(xm, 0, wn, lane, red, mrow, irow);  // its barriers also retire every read of O
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int n = wn + j * 16 + r16;
            const float g = a.n2g[n], b = a.n2b[n];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = i * 16 + rq + r;
                    put_a<SPLIT>(Hh, Hl, row, n, (xm[i][j][r] - mrow[row]) * irow[row] * g + b);
                }
        }
        __syncthreads();
    }

    // ---- MLP: 128 intermediate columns at a time
    f4v acc2[4][J];
    zero(acc2);
    load_w(w1, a.W1, RD, 0, 0, wn, r16, kq);
    for (int c0 = 0; c0 < ff; c0 += RD) {
        load_w(w2, a.W2, ff, 0, c0, wn, r16, kq);  // in flight during ff1
        f4v acc1[4][J];
        zero(acc1);
        mma_panel<SPLIT, J>(acc1, Hh, Hl, w1, r16, kq);
        float bb[J];
#pragma unroll
        for (int j = 0; j < J; j++) bb[j] = a.b1[c0 + wn + j * 16 + r16];
        if (c0 + RD < ff) load_w(w1, a.W1, RD, c0 + RD, 0, wn, r16, kq);  // in flight during GELU and ff2
        __syncthreads();  // every wave has finished reading the previous intermediate chunk
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < J; j++)
#pragma unroll
                for (int r = 0; r < 4; r++) put_a<SPLIT>(Fh, Fl, i * 16 + rq + r, wn + j * 16 + r16, gelu_fast(acc1[i][j][r] + bb[j]));
        __syncthreads();
        mma_panel<SPLIT, J>(acc2, Fh, Fl, w2, r16, kq);
    }
#pragma unroll
    for (int j = 0; j < J; j++) {
        const float b = a.b2[wn + j * 16 + r16];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + i * 16 + rq + r;
                const float v = xm[i][j][r] + acc2[i][j][r] + b;
                xm[i][j][r] = v;
                if (m < M) a.X[(long)m * RD + wn + j * 16 + r16] = v;
            }
    }
    if (!QKV && !LNOUT) return;

    // ---- the next block's LN1, then its qkv product or H itself
    ln_row_stats<RB, J>(xm, 0, wn, lane, red, mrow, irow);  // its barriers also retire every read of H
    if (QKV) load_w(w1, a.Wqkv, RD, 0, 0, wn, r16, kq);  // in flight while H is written
#pragma unroll
    for (int j = 0; j < J; j++) {
        const int n = wn + j * 16 + r16;
        const float g = a.n1g[n], b = a.n1b[n];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float h = (xm[i][j][r] - mrow[i * 16 + rq + r]) * irow[i * 16 + rq + r] * g + b;
                if (QKV) {
                    put_a<SPLIT>(Hh, Hl, i * 16 + rq + r, n, h);
                } else {
                    const int m = m0 + i * 16 + rq + r;
                    if (m >= M) continue;
                    if (SPLIT) static_cast<float*>(a.Hout)[(long)m * RD + n] = h;
                    else static_cast<uint16_t*>(a.Hout)[(long)m * RD + n] = f2bf(h);
                }
            }
    }
    if (!QKV) return;
    __syncthreads();
#pragma unroll 1
    for (int c = 0; c < 3; c++) {
        f4v acc[4][J];
        zero(acc);
        mma_panel<SPLIT, J>(acc, Hh, Hl, w1, r16, kq);
        if (c < 2) load_w(w1, a.Wqkv, RD, (c + 1) * RD, 0, wn, r16, kq);
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int n = c * RD + wn + j * 16 + r16;
            const float b = a.bqkv[n];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int m = m0 + i * 16 + rq + r;
                    if (m >= M) continue;
                    const float v = acc[i][j][r] + b;
                    if (SPLIT) static_cast<float*>(a.QKV)[(long)m * 3 * RD + n] = v;
                    else static_cast<uint16_t*>(a.QKV)[(long)m * 3 * RD + n] = f2bf(v);
                }
        }
    }
}

}  // namespace

bool rows_fused_ok(int dim, int ff) { return dim == RD && ff % RD == 0; }

void rows_fused(hipStream_t s, int M, int dim, int ff, bool split, const RowsArgs& a, int waves) {
    if (M <= 0) return;
    if (!rows_fused_ok(dim, ff)) fail(J2K_ERR_LIMIT, "fused row kernel needs dim == 128 and ff a multiple of 128");
    const bool proj = a.Wp != nullptr, qkv = a.Wqkv != nullptr, lnout = !qkv && a.n1g != nullptr;
    if (proj ? !a.O : !a.Hin) fail(J2K_ERR_INVALID, "fused row kernel: missing input");
    if (waves != 4 && waves != 8) fail(J2K_ERR_INVALID, "fused row kernel: 4 or 8 waves");
    dim3 grid(cdiv(M, RB)), block(64 * waves);
#define J2K_ROWS(S, P, Q, L)                                                                   \
    if (split == S && proj == P && qkv == Q && lnout == L) {                                   \
        if (waves == 4) hipLaunchKernelGGL((k_rows<S, P, Q, L, 4>), grid, block, 0, s, M, ff, a); \
        else hipLaunchKernelGGL((k_rows<S, P, Q, L, 8>), grid, block, 0, s, M, ff, a);           \
        LAUNCHCHK("rows_fused");                                                               \
        return;                                                                                \
    }
    J2K_ROWS(true, true, true, false)
    J2K_ROWS(true, true, false, false)
    J2K_ROWS(true, false, false, true)
    J2K_ROWS(true, false, false, false)
    J2K_ROWS(false, true, true, false)
    J2K_ROWS(false, true, false, false)
    J2K_ROWS(false, false, false, true)
    J2K_ROWS(false, false, false, false)
#undef J2K_ROWS
    fail(J2K_ERR_INVALID, "fused row kernel: unsupported stage combination");
}

}  // namespace j2k
