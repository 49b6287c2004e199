// Row-panel kernel for the d = 128 trunk: everything of a block
// (cpu.rs:287-313) that is local to a token, fused over 64-token panels so
// the panel's activations stay in LDS and registers:
//
//   PROJ   X += O @ Wp + bp, then H = LN2(X)         (else H comes from Hin)
//          X += ff2(GELU(ff1(H)))                    (always)
//   QKV    QKV = LN1'(X) @ Wqkv' + bqkv'             (next block's norm and qkv)
//   LNOUT  Hout = LN1'(X)                            (when the qkv GEMM runs apart)
//   RO     the cell_st readout of X (last block)     (X itself not stored)
//
// With PROJ and QKV a layer is two launches: this kernel and the attention.
// The same arithmetic, in the same order, as the tile GEMM epilogues, so
// every combination gives bit-identical results and the autotuner may pick
// any of them.
//
// A block owns 64 tokens; in the main launch shape its four waves split
// every 128-column product (32 columns each, 2 x 4 tiles of
// v_mfma_f32_16x16x32_bf16, f32 accumulate; the other shapes below).
// The products run transposed: the MFMA's first operand is a 16 x 32
// fragment of W (K-contiguous bf16 [N, K], so output columns are its rows),
// the second a 32 x 16 fragment of the panel (tokens) from LDS, as bf16 or,
// SPLIT, as a bf16 hi + lo pair of an f32 value (two MFMAs per tile). Each
// output element is the same dot product over K as in the untransposed
// tile GEMM. A lane's accumulator then holds four consecutive columns of
// one token: residual, biases and results move as 16-byte vectors (raw
// buffer operations bounded by the panel's rows, so a ragged last panel
// needs no branches), and the next product's operand is written as 8-byte
// LDS stores. Weights (in fragment order, 1 KB per wave load) stream from
// L2 a whole chunk ahead of their use. The ff-wide intermediate is made 128
// (or 64) columns at a time and consumed at once as the next K chunk of ff2.
// The residual after the projection stays in registers through the MLP in
// the main shape; the others park it in X and read it back. Row statistics
// of the LayerNorms in the order every served kernel uses (lnstats.h).
#include "common.h"
#include "kernels.h"
#include "embedln.h"
#include "lnstats.h"

namespace j2k {

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));
typedef unsigned u2v __attribute__((ext_vector_type(2)));

constexpr int RB = 64, RD = 128;

__device__ __forceinline__ bf8v frag(const uint16_t* p) { return __builtin_bit_cast(bf8v, *reinterpret_cast<const uint4*>(p)); }

// An LDS panel (64 token rows of bf16, any number of columns) is stored as
// 16-byte chunks of 8 columns, chunk-major, the 64 rows of chunk ch permuted
// by row ^ (ch & 7): element (row, col) sits at pidx(row, col). An MFMA
// operand read (16 lanes, 16 rows of one chunk) is then one 256-byte span
// (conflict-free on gfx950's 64 banks; row-major rows of 272 bytes put the
// 8 rows of each 8-lane half on 8 lines and take two passes), and the eight
// chunks of one row that a staging store writes land on distinct banks. A
// wave's first row (a multiple of 32) is a plain offset of wm * 8.
__device__ __forceinline__ int pidx(int row, int col) {
    const int ch = col >> 3;
    return ((ch * RB + (row ^ (ch & 7))) << 3) | (col & 7);
}

// A raw buffer over `bytes` from p: loads past the end read 0, stores past it
// are dropped.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rows_rsrc(const void* p, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, bytes, 0x00020000);
}
__device__ __forceinline__ f4v bload4(__amdgpu_buffer_rsrc_t r, int byte_off) {
    return __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, 0));
}
__device__ __forceinline__ void bstore4(__amdgpu_buffer_rsrc_t r, int byte_off, f4v v) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4v, v), r, byte_off, 0, 0);
}
// Four bf16 (round to nearest even) of v, first element at the lower address.
__device__ __forceinline__ u2v bf4(f4v v) { return u2v{f2bf2(v[0], v[1]), f2bf2(v[2], v[3])}; }
__device__ __forceinline__ void bstore4(__amdgpu_buffer_rsrc_t r, int byte_off, u2v v) {
    __builtin_amdgcn_raw_buffer_store_b64(v, r, byte_off, 0, 0);
}

// Fragments of W^T rows n0 + wn + j*16 .. +16 (j < J), K columns k0 .. k0+32*K4,
// from W in fragment order (16-row, 32-column blocks of 64 lanes x 8 values,
// K-blocks fastest; ldk = K): every load is 1 KB in a row.
template <int J, int K4 = 4>
__device__ __forceinline__ void load_w(bf8v (&w)[K4][J], const uint16_t* Wf, int ldk, int n0, int k0, int wn, int lane) {
    const uint16_t* p = Wf + (((long)(n0 + wn) / 16 * (ldk / 32) + k0 / 32) * 64 + lane) * 8;
#pragma unroll
    for (int k4 = 0; k4 < K4; k4++)
#pragma unroll
        for (int j = 0; j < J; j++) w[k4][j] = frag(p + ((long)j * (ldk / 32) + k4) * 512);
}

// acc += (A (16*MI x 32*K4, an LDS panel, hi and lo) @ the loaded W chunk),
// transposed: acc[i][j][r] is A's row i*16 + r16, column
// wn + j*16 + 4*(lane/16) + r.
template <bool SPLIT, int MI, int J, int K4 = 4>
__device__ __forceinline__ void mma_panel(f4v (&acc)[MI][J], const uint16_t* Ah, const uint16_t* Al,
                                          const bf8v (&w)[K4][J], int r16, int kq) {
    // pidx(i*16 + r16, k4*32 + kq) is a lane base (one for even k4, one for
    // odd: the row permutation of chunk k4*4 + kq/8) plus a constant
    // (odd k4 flips row bit 2, element bit 5)
    const int sq = kq >> 3;
    const int b0 = (sq * RB + (r16 ^ sq)) << 3;
    auto at = [&](int k4, int i) { return (k4 & 1 ? b0 ^ 32 : b0) + ((k4 * 4 * RB + i * 16) << 3); };
    // the operands of step k4 + 1 are read while step k4 multiplies
    bf8v a[2][MI], l[2][MI];
#pragma unroll
    for (int i = 0; i < MI; i++) {
        a[0][i] = frag(&Ah[at(0, i)]);
        if (SPLIT) l[0][i] = frag(&Al[at(0, i)]);
    }
    __builtin_amdgcn_s_setprio(1);  // the other wave on the SIMD fills the MFMA gaps
#pragma unroll
    for (int k4 = 0; k4 < K4; k4++) {
        const int b = k4 & 1;
        if (k4 + 1 < K4)
#pragma unroll
            for (int i = 0; i < MI; i++) {
                a[b ^ 1][i] = frag(&Ah[at(k4 + 1, i)]);
                if (SPLIT) l[b ^ 1][i] = frag(&Al[at(k4 + 1, i)]);
            }
#pragma unroll
        for (int i = 0; i < MI; i++)
#pragma unroll
            for (int j = 0; j < J; j++) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[k4][j], a[b][i], acc[i][j], 0, 0, 0);
                if (SPLIT) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[k4][j], l[b][i], acc[i][j], 0, 0, 0);
            }
    }
    __builtin_amdgcn_s_setprio(0);
}

template <int MI, int J>
__device__ __forceinline__ void zero(f4v (&acc)[MI][J]) {
#pragma unroll
    for (int i = 0; i < MI; i++)
#pragma unroll
        for (int j = 0; j < J; j++) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};
}

// Four consecutive columns (col a multiple of 4) of one panel row as A
// operands: bf16, and with SPLIT the bf16 remainder.
template <bool SPLIT>
__device__ __forceinline__ void put_a4(uint16_t* Ah, uint16_t* Al, int row, int col, f4v v) {
    const u2v h = bf4(v);
    const int o = pidx(row, col);
    *reinterpret_cast<u2v*>(&Ah[o]) = h;
    if (SPLIT) {
        const f4v r = {v[0] - __uint_as_float(h[0] << 16), v[1] - __uint_as_float(h[0] & 0xffff0000u),
                       v[2] - __uint_as_float(h[1] << 16), v[3] - __uint_as_float(h[1] & 0xffff0000u)};
        *reinterpret_cast<u2v*>(&Al[o]) = bf4(r);
    }
}

// The panel's rows x 128 of global A (bf16, or f32 split into hi + lo) into
// LDS; rows past the panel's end read as 0.
template <bool SPLIT, int NT>
__device__ __forceinline__ void stage_rows(const void* src, int m0, int rows, uint16_t* Ah, uint16_t* Al, int tid) {
    if (!SPLIT) {
        const auto r = rows_rsrc(static_cast<const uint16_t*>(src) + (long)m0 * RD, rows * RD * 2);
#pragma unroll
        for (int q = 0; q < 1024 / NT; q++) {
            const int v = tid + NT * q, row = v >> 4, kv = v & 15;
            const u4v val = __builtin_amdgcn_raw_buffer_load_b128(r, (row * RD + kv * 8) * 2, 0, 0);
            *reinterpret_cast<u4v*>(&Ah[pidx(row, kv * 8)]) = val;
        }
    } else {
        const auto r = rows_rsrc(static_cast<const float*>(src) + (long)m0 * RD, rows * RD * 4);
#pragma unroll
        for (int q = 0; q < 2048 / NT; q++) {
            const int v = tid + NT * q, row = v >> 5, kv = v & 31;
            put_a4<true>(Ah, Al, row, kv * 4, bload4(r, (row * RD + kv * 4) * 4));
        }
    }
}

// The cell_st embedding of tile rows m0 + i*16 + r16 (i < 4), columns
// wn + 16j + c4 .. +4 (j < J): bias, then the four features in order
// (zero features skipped, as cpu.rs:26-45), then the request's meta vector.
// Rows past M repeat row M - 1 (never stored).
template <int J>
__device__ __forceinline__ void embed_rows(f4v (&xv)[4][J], const RowsEmbed& e, int M, int m0, int wn, int r16,
                                           int c4) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int m = m0 + i * 16 + r16, mm = m < M ? m : M - 1;
        const float4 f = reinterpret_cast<const float4*>(e.x4)[mm];
        const float* mv = e.mvec + (e.tok_req ? (long)e.tok_req[mm] * RD : 0);
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int d0 = wn + j * 16 + c4;
            const f4v bv = *reinterpret_cast<const f4v*>(e.be + d0);
            const f4v w0 = *reinterpret_cast<const f4v*>(e.we + d0);
            const f4v w1 = *reinterpret_cast<const f4v*>(e.we + RD + d0);
            const f4v w2 = *reinterpret_cast<const f4v*>(e.we + 2 * RD + d0);
            const f4v w3 = *reinterpret_cast<const f4v*>(e.we + 3 * RD + d0);
            const f4v mw = *reinterpret_cast<const f4v*>(mv + d0);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float v = bv[k];
                if (f.x != 0.f) v += f.x * w0[k];
                if (f.y != 0.f) v += f.y * w1[k];
                if (f.z != 0.f) v += f.z * w2[k];
                if (f.w != 0.f) v += f.w * w3[k];
                v += mw[k];
                xv[i][j][k] = v;
            }
        }
    }
}

// The readout of the block's tokens whose final residual xm is held as in
// the row-panel kernel (row wm + 16i + r16, columns wn + 16j + 4(lane/16) + r):
// LN with the statistics' tree, then each of the three outputs as a row
// total of the products h_d w[d][c] in the same tree, plus the bias. scratch:
// 8 KB of idle LDS; ys: [3][RB] LDS.
template <int MI, int J>
__device__ __forceinline__ void readout_rows(const f4v (&xm)[MI][J], int wm, int wn, int lane, const Readout& ro,
                                             int m0, int rows, void* scratch, float* mrow, float* irow,
                                             float (*ys)[RB]) {
#pragma clang fp contract(off)
    float(*red)[RB][4] = reinterpret_cast<float(*)[RB][4]>(scratch);
    const int c4 = (lane >> 4) * 4, r16 = lane & 15;
    ln_row_stats_t<RB, MI, J>(xm, wm, wn, lane, red, mrow, irow);
    f4v h[MI][J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        const f4v g = *reinterpret_cast<const f4v*>(ro.g + wn + j * 16 + c4);
        const f4v b = *reinterpret_cast<const f4v*>(ro.b + wn + j * 16 + c4);
#pragma unroll
        for (int i = 0; i < MI; i++) {
            const int row = wm + i * 16 + r16;
            const float mu = mrow[row], iv = irow[row];
#pragma unroll
            for (int r = 0; r < 4; r++) h[i][j][r] = (xm[i][j][r] - mu) * iv * g[r] + b[r];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        f4v p[MI][J];
#pragma unroll
        for (int j = 0; j < J; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float w = ro.w[(wn + j * 16 + c4 + r) * 3 + c];
#pragma unroll
                for (int i = 0; i < MI; i++) p[i][j][r] = h[i][j][r] * w;
            }
        row_sums_t<RB, MI, J>(p, wm, wn, lane, red, ys[c]);
    }
    const int t = threadIdx.x;
    if (t < rows) {
        const long m = (long)m0 + t;
        const float y0 = ys[0][t] + ro.bias[0], y1 = ys[1][t] + ro.bias[1], y2 = ys[2][t] + ro.bias[2];
        if (ro.y) {
            ro.y[m * 3] = y0;
            ro.y[m * 3 + 1] = y1;
            ro.y[m * 3 + 2] = y2;
        }
        if (ro.xn) {
            const float4 a = reinterpret_cast<const float4*>(ro.x0)[m];
            reinterpret_cast<float4*>(ro.xn)[m] = a.w != 0.f ? a : make_float4(a.x + y0, a.y + y1, a.z + y2, 0.f);
        }
    }
}

// The block's waves form a WM x WN grid over (tokens, product columns): a
// wave owns 64/WM tokens and 128/WN columns of every 128-column product.
// WM = 1, WN = 4 (32 columns a wave) reads the whole 64-token panel from LDS
// for every product; WM = WN = 2 halves those reads per MFMA and loads twice
// the weight fragments; WN = 8 (16 columns) holds half the registers, with
// twice the waves per block. FC: intermediate columns per MLP chunk; at 64
// the block's LDS (48.8 KB in the split form) lets three blocks share a CU.
// The arithmetic of an output element is the same in every shape. The
// LayerNorm row sums borrow LDS that is idle at the time: F's before the
// MLP, H's after it.
template <bool SPLIT, bool PROJ, bool QKV, bool LNOUT, int WM, int WN, int FC, bool RO = false, bool EMB = false>
__global__ __launch_bounds__(64 * WM * WN, (WM * WN == 4 && FC == 64) ? 3 : WM * WN / 2) void k_rows(int M, int ff,
                                                                                                     RowsArgs a) {
    static_assert(!RO || (PROJ && !QKV && !LNOUT), "the readout ends the last block");
    static_assert(!EMB || (PROJ && QKV && WM == 1 && WN == 4), "the embedding enters the first block");
    constexpr int WAVES = WM * WN, MI = 4 / WM, J = 8 / WN, NT = 64 * WAVES, P = SPLIT ? 2 : 1;
    constexpr int JF = FC / (16 * WN), K4F = FC / 32;  // ff1 tiles, ff2 K steps
    static_assert(JF >= 1 && FC % 32 == 0, "each wave owns at least one 16-column tile of an MLP chunk");
    static_assert(RB * FC * P * 2 >= 8 * RB * 16, "the LayerNorm row sums fit in F");
    __shared__ __align__(16) uint16_t Hs[P * RB * RD];
    __shared__ __align__(16) uint16_t Fs[P * RB * FC];
    __shared__ float mrow[RB], irow[RB];
    __shared__ float ys[RO ? 3 : 1][RB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave / WN) * 16 * MI, wnx = wave % WN;  // the wave's first token, its column slot
    const int m0 = blockIdx.x * RB, wn = wnx * 16 * J, wf = wnx * (FC / WN), rows = min(RB, M - m0);
    const int r16 = lane & 15, kq = (lane >> 4) * 8, c4 = (lane >> 4) * 4;
    // the wave's tokens of H and F (split: hi, lo)
    uint16_t *const Hh = Hs + wm * 8, *const Hl = Hs + (SPLIT ? RB * RD : 0) + wm * 8;
    uint16_t *const Fh = Fs + wm * 8, *const Fl = Fs + (SPLIT ? RB * FC : 0) + wm * 8;
    const auto rX = rows_rsrc(a.X + (long)m0 * RD, rows * RD * 4);
    // byte offset of the lane's four columns of tile (i, j) in a [64][ld] f32 panel
    auto at = [&](int i, int j, int ld) { return ((wm + i * 16 + r16) * ld + wn + j * 16 + c4) * 4; };
    bf8v w1[4][J];

    // With the register budget of the four-wave 128-column shape the
    // projected residual stays in registers through the MLP; the other
    // shapes park it in X and read it back.
    constexpr bool KEEP = PROJ && WM == 1 && WN == 4 && FC == 128;
    f4v xk[KEEP ? MI : 1][J];

    // ---- H, the MLP's input: LN2 of the projected residual, or Hin
    stage_rows<SPLIT, NT>(PROJ ? a.O : a.Hin, m0, rows, Hs, Hs + (SPLIT ? RB * RD : 0), tid);
    if (PROJ) load_w(w1, a.Wp, RD, 0, 0, wn, lane);
    __syncthreads();
    if (PROJ) {
        f4v xm[MI][J];
        if constexpr (EMB) {
            embed_rows<J>(xm, a.emb, M, m0, wn, r16, c4);
        } else {
#pragma unroll
            for (int i = 0; i < MI; i++)
#pragma unroll
                for (int j = 0; j < J; j++) xm[i][j] = bload4(rX, at(i, j, RD));
        }
        f4v acc[MI][J];
        zero(acc);
        mma_panel<SPLIT, MI, J>(acc, Hh, Hl, w1, r16, kq);
#pragma unroll
        for (int j = 0; j < J; j++) {
            const f4v b = *reinterpret_cast<const f4v*>(a.bp + wn + j * 16 + c4);
#pragma unroll
            for (int i = 0; i < MI; i++) {
#pragma unroll
                for (int r = 0; r < 4; r++) xm[i][j][r] = xm[i][j][r] + acc[i][j][r] + b[r];
                if constexpr (KEEP) xk[i][j] = xm[i][j];
                else bstore4(rX, at(i, j, RD), xm[i][j]);  // read back after the MLP
            }
        }
        // its barriers also retire every read of O
        ln_row_stats_t<RB, MI, J>(xm, wm, wn, lane, reinterpret_cast<float(*)[RB][4]>(Fs), mrow, irow);
#pragma unroll
        for (int j = 0; j < J; j++) {
            const f4v g = *reinterpret_cast<const f4v*>(a.n2g + wn + j * 16 + c4);
            const f4v b = *reinterpret_cast<const f4v*>(a.n2b + wn + j * 16 + c4);
#pragma unroll
            for (int i = 0; i < MI; i++) {
                const int row = wm + i * 16 + r16;
                const float mu = mrow[row], iv = irow[row];
                f4v h;
#pragma unroll
                for (int r = 0; r < 4; r++) h[r] = (xm[i][j][r] - mu) * iv * g[r] + b[r];
                put_a4<SPLIT>(Hh, Hl, i * 16 + r16, wn + j * 16 + c4, h);
            }
        }
        __syncthreads();
    }

    // ---- MLP: FC intermediate columns at a time
    f4v acc2[MI][J];
    zero(acc2);
    bf8v v1[4][JF], w2[K4F][J];
    load_w<JF>(v1, a.W1, RD, 0, 0, wf, lane);
    for (int c0 = 0; c0 < ff; c0 += FC) {
        load_w<J, K4F>(w2, a.W2, ff, 0, c0, wn, lane);  // in flight during ff1
        f4v acc1[MI][JF];
        zero(acc1);
        mma_panel<SPLIT, MI, JF>(acc1, Hh, Hl, v1, r16, kq);
        f4v bb[JF];
#pragma unroll
        for (int j = 0; j < JF; j++) bb[j] = *reinterpret_cast<const f4v*>(a.b1 + c0 + wf + j * 16 + c4);
        if (c0 + FC < ff) load_w<JF>(v1, a.W1, RD, c0 + FC, 0, wf, lane);  // in flight during GELU and ff2
        __syncthreads();  // every wave has finished reading the previous intermediate chunk
#pragma unroll
        for (int i = 0; i < MI; i++)
#pragma unroll
            for (int j = 0; j < JF; j++) {
                f4v v;
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = gelu_fast(acc1[i][j][r] + bb[j][r]);
                put_a4<SPLIT>(Fh, Fl, i * 16 + r16, wf + j * 16 + c4, v);
            }
        __syncthreads();
        mma_panel<SPLIT, MI, J, K4F>(acc2, Fh, Fl, w2, r16, kq);
    }
    f4v xm[MI][J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        const f4v b = *reinterpret_cast<const f4v*>(a.b2 + wn + j * 16 + c4);
#pragma unroll
        for (int i = 0; i < MI; i++) {
            if constexpr (KEEP) xm[i][j] = xk[i][j];
            else xm[i][j] = bload4(rX, at(i, j, RD));
#pragma unroll
            for (int r = 0; r < 4; r++) xm[i][j][r] = xm[i][j][r] + acc2[i][j][r] + b[r];
            if constexpr (!RO) bstore4(rX, at(i, j, RD), xm[i][j]);
        }
    }
    if constexpr (RO) {
        // every ff1 read of H is done: its LDS holds the sums
        readout_rows<MI, J>(xm, wm, wn, lane, a.ro, m0, rows, Hs, mrow, irow, ys);
        return;
    }
    if (!QKV && !LNOUT) return;

    // ---- the next block's LN1, then its qkv product or H itself
    // (every ff1 read of H is done)
    ln_row_stats_t<RB, MI, J>(xm, wm, wn, lane, reinterpret_cast<float(*)[RB][4]>(Hs), mrow, irow);
    if (QKV) load_w(w1, a.Wqkv, RD, 0, 0, wn, lane);  // in flight while H is written
    const auto rH = rows_rsrc(LNOUT ? static_cast<const char*>(a.Hout) + (long)m0 * RD * (SPLIT ? 4 : 2) : nullptr,
                              LNOUT ? rows * RD * (SPLIT ? 4 : 2) : 0);
#pragma unroll
    for (int j = 0; j < J; j++) {
        const f4v g = *reinterpret_cast<const f4v*>(a.n1g + wn + j * 16 + c4);
        const f4v b = *reinterpret_cast<const f4v*>(a.n1b + wn + j * 16 + c4);
#pragma unroll
        for (int i = 0; i < MI; i++) {
            const int row = wm + i * 16 + r16;
            const float mu = mrow[row], iv = irow[row];
            f4v h;
#pragma unroll
            for (int r = 0; r < 4; r++) h[r] = (xm[i][j][r] - mu) * iv * g[r] + b[r];
            if (QKV) put_a4<SPLIT>(Hh, Hl, i * 16 + r16, wn + j * 16 + c4, h);
            else if (SPLIT) bstore4(rH, at(i, j, RD), h);
            else bstore4(rH, at(i, j, RD) / 2, bf4(h));
        }
    }
    if (!QKV) return;
    __syncthreads();
    const auto rQ = rows_rsrc(static_cast<const char*>(a.QKV) + (long)m0 * 3 * RD * (SPLIT ? 4 : 2),
                              rows * 3 * RD * (SPLIT ? 4 : 2));
#pragma unroll 1
    for (int c = 0; c < 3; c++) {
        // Under the smaller register budgets, H's fragments are read again
        // for each chunk rather than held across all three (which spills).
        if (WAVES == 8 || FC == 64 || WM == 2) asm volatile("" ::: "memory");
        f4v acc[MI][J];
        zero(acc);
        mma_panel<SPLIT, MI, J>(acc, Hh, Hl, w1, r16, kq);
        if (c < 2) load_w(w1, a.Wqkv, RD, (c + 1) * RD, 0, wn, lane);
#pragma unroll
        for (int j = 0; j < J; j++) {
            const f4v b = *reinterpret_cast<const f4v*>(a.bqkv + c * RD + wn + j * 16 + c4);
#pragma unroll
            for (int i = 0; i < MI; i++) {
                f4v v;
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = acc[i][j][r] + b[r];
                const int o = at(i, j, 3 * RD) + c * RD * 4;
                if (SPLIT) bstore4(rQ, o, v);
                else bstore4(rQ, o / 2, bf4(v));
            }
        }
    }
}

// The cell_st embedding, the first block's LN1 and its qkv product in one
// launch (d = 128; the served forms): X = embed(x4) + bias + meta vector
// (the per-element arithmetic of k_cell_embed_ln), LN1 with the row-panel
// kernel's statistics, H in LDS only, QKV = H @ Wqkv + bqkv as in
// k_rows_qkv. 64 tokens a block; four waves of 32 columns. Block 0 also
// clears a.zero[0 .. a.n_zero).
template <bool SPLIT>
__global__ __launch_bounds__(256, 2) void k_rows_embed_qkv(int M, const float4* __restrict__ x4,
                                                           const float* __restrict__ we, const float* __restrict__ be,
                                                           const float* __restrict__ mvec,
                                                           const int* __restrict__ tok_req, RowsArgs a) {
    constexpr int J = 2, P = SPLIT ? 2 : 1;
    __shared__ __align__(16) uint16_t Hs[P * RB * RD];
    __shared__ float mrow[RB], irow[RB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * RB, wn = wave * 32, rows = min(RB, M - m0);
    const int r16 = lane & 15, kq = (lane >> 4) * 8, c4 = (lane >> 4) * 4;
    uint16_t *const Hh = Hs, *const Hl = Hs + (SPLIT ? RB * RD : 0);
    bf8v w[4][J];
    load_w(w, a.Wqkv, RD, 0, 0, wn, lane);
    if (blockIdx.x == 0)
        for (int i = tid; i < a.n_zero; i += 256) a.zero[i] = 0u;
    // no X: a zero-sized buffer drops the stores
    const auto rX = rows_rsrc(a.X ? a.X + (long)m0 * RD : nullptr, a.X ? rows * RD * 4 : 0);
    f4v xv[4][J];
    RowsEmbed e;
    e.x4 = reinterpret_cast<const float*>(x4);
    e.we = we;
    e.be = be;
    e.mvec = mvec;
    e.tok_req = tok_req;
    embed_rows<J>(xv, e, M, m0, wn, r16, c4);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < J; j++) bstore4(rX, ((i * 16 + r16) * RD + wn + j * 16 + c4) * 4, xv[i][j]);
    ln_row_stats_t<RB, 4, J>(xv, 0, wn, lane, reinterpret_cast<float(*)[RB][4]>(Hs), mrow, irow);
#pragma unroll
    for (int j = 0; j < J; j++) {
        const f4v g = *reinterpret_cast<const f4v*>(a.n1g + wn + j * 16 + c4);
        const f4v b = *reinterpret_cast<const f4v*>(a.n1b + wn + j * 16 + c4);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = i * 16 + r16;
            const float mu = mrow[row], iv = irow[row];
            f4v h;
#pragma unroll
            for (int r = 0; r < 4; r++) h[r] = (xv[i][j][r] - mu) * iv * g[r] + b[r];
            put_a4<SPLIT>(Hh, Hl, row, wn + j * 16 + c4, h);
        }
    }
    __syncthreads();
    const auto rQ = rows_rsrc(static_cast<const char*>(a.QKV) + (long)m0 * 3 * RD * (SPLIT ? 4 : 2),
                              rows * 3 * RD * (SPLIT ? 4 : 2));
#pragma unroll 1
    for (int c = 0; c < 3; c++) {
        asm volatile("" ::: "memory");  // H's fragments read per chunk, not held across all three
        f4v acc[4][J];
        zero(acc);
        mma_panel<SPLIT, 4, J>(acc, Hh, Hl, w, r16, kq);
        if (c < 2) load_w(w, a.Wqkv, RD, (c + 1) * RD, 0, wn, lane);
#pragma unroll
        for (int j = 0; j < J; j++) {
            const f4v b = *reinterpret_cast<const f4v*>(a.bqkv + c * RD + wn + j * 16 + c4);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                f4v v;
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = acc[i][j][r] + b[r];
                const int o = ((i * 16 + r16) * 3 * RD + c * RD + wn + j * 16 + c4) * 4;
                if (SPLIT) bstore4(rQ, o, v);
                else bstore4(rQ, o / 2, bf4(v));
            }
        }
    }
}

// The readout alone (X [M][128] in memory), with readout_rows' arithmetic.
__global__ __launch_bounds__(256, 2) void k_rows_readout(int M, const float* __restrict__ X, Readout ro) {
    constexpr int J = 2;
    __shared__ __align__(16) float scratch[8 * RB * 4];
    __shared__ float mrow[RB], irow[RB], ys[3][RB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * RB, wn = wave * 32, rows = min(RB, M - m0);
    const int r16 = lane & 15, c4 = (lane >> 4) * 4;
    const auto rX = rows_rsrc(X + (long)m0 * RD, rows * RD * 4);
    f4v xm[4][J];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < J; j++) xm[i][j] = bload4(rX, ((i * 16 + r16) * RD + wn + j * 16 + c4) * 4);
    readout_rows<4, J>(xm, 0, wn, lane, ro, m0, rows, scratch, mrow, irow, ys);
}

// The first layer's qkv product on its own (the embedding wrote H): the
// row-panel kernel's last stage, QKV = H @ Wqkv + bqkv for 64 tokens a block,
// with the same arithmetic.
template <bool SPLIT>
__global__ __launch_bounds__(256, 2) void k_rows_qkv(int M, RowsArgs a) {
    constexpr int J = 2, P = SPLIT ? 2 : 1;
    __shared__ __align__(16) uint16_t Hs[P * RB * RD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * RB, wn = wave * 32, rows = min(RB, M - m0);
    const int r16 = lane & 15, kq = (lane >> 4) * 8, c4 = (lane >> 4) * 4;
    uint16_t *const Hh = Hs, *const Hl = Hs + (SPLIT ? RB * RD : 0);
    bf8v w[4][J];
    load_w(w, a.Wqkv, RD, 0, 0, wn, lane);
    stage_rows<SPLIT, 256>(a.Hin, m0, rows, Hh, Hl, tid);
    __syncthreads();
    const auto rQ = rows_rsrc(static_cast<const char*>(a.QKV) + (long)m0 * 3 * RD * (SPLIT ? 4 : 2),
                              rows * 3 * RD * (SPLIT ? 4 : 2));
#pragma unroll 1
    for (int c = 0; c < 3; c++) {
        f4v acc[4][J];
        zero(acc);
        mma_panel<SPLIT, 4, J>(acc, Hh, Hl, w, r16, kq);
        if (c < 2) load_w(w, a.Wqkv, RD, (c + 1) * RD, 0, wn, lane);
#pragma unroll
        for (int j = 0; j < J; j++) {
            const f4v b = *reinterpret_cast<const f4v*>(a.bqkv + c * RD + wn + j * 16 + c4);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                f4v v;
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = acc[i][j][r] + b[r];
                const int o = ((i * 16 + r16) * 3 * RD + c * RD + wn + j * 16 + c4) * 4;
                if (SPLIT) bstore4(rQ, o, v);
                else bstore4(rQ, o / 2, bf4(v));
            }
        }
    }
}

// The MLP of a wider trunk (D = 256: the wide shape) over 64-token panels,
// X += ff2(GELU(ff1(H))) + biases: the ff-wide intermediate never reaches
// HBM (at 262,144 tokens and ff 1024 it is 1 GB of f32 per block otherwise,
// written by one GEMM and read back by the next). Eight waves, each owning
// 32 of the D output columns of ff2 and 16 of the 128 columns of an ff1
// chunk, over all 64 tokens; H (hi + lo in the split form) and the current
// intermediate chunk sit in LDS (96 KB split: one block per CU). Products
// as in k_rows (transposed, weights in fragment order from L2 one chunk
// ahead), so every output is the same dot product, in the same order, with
// the same bias, GELU and residual arithmetic as the tile GEMMs' GELU and
// residual epilogues: the two arrangements give bit-identical results.
template <bool SPLIT, int D>
__global__ __launch_bounds__(512, 1) void k_mlp_rows(int M, int ff, RowsArgs a) {
    constexpr int WN = 8, NT = 64 * WN, MI = 4, J = D / (16 * WN), FC = 128, JF = FC / (16 * WN);
    constexpr int K4D = D / 32, K4F = FC / 32, P = SPLIT ? 2 : 1;
    static_assert(J >= 1 && JF == 1, "eight waves over D output columns and a 128-column chunk");
    __shared__ __align__(16) uint16_t Hs[P * RB * D];
    __shared__ __align__(16) uint16_t Fs[P * RB * FC];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * RB, rows = min(RB, M - m0);
    const int wn = wave * 16 * J, wf = wave * 16;
    const int r16 = lane & 15, kq = (lane >> 4) * 8, c4 = (lane >> 4) * 4;
    uint16_t *const Hh = Hs, *const Hl = Hs + (SPLIT ? RB * D : 0);
    uint16_t *const Fh = Fs, *const Fl = Fs + (SPLIT ? RB * FC : 0);
    // H: the panel's rows x D (bf16, or f32 split into hi + lo); rows past the end read 0
    if (SPLIT) {
        const auto r = rows_rsrc(static_cast<const float*>(a.Hin) + (long)m0 * D, rows * D * 4);
#pragma unroll
        for (int q = 0; q < RB * D / 4 / NT; q++) {
            const int v = tid + NT * q, row = v / (D / 4), kv = v % (D / 4);
            put_a4<true>(Hh, Hl, row, kv * 4, bload4(r, (row * D + kv * 4) * 4));
        }
    } else {
        const auto r = rows_rsrc(static_cast<const uint16_t*>(a.Hin) + (long)m0 * D, rows * D * 2);
#pragma unroll
        for (int q = 0; q < RB * D / 8 / NT; q++) {
            const int v = tid + NT * q, row = v / (D / 8), kv = v % (D / 8);
            const u4v val = __builtin_amdgcn_raw_buffer_load_b128(r, (row * D + kv * 8) * 2, 0, 0);
            *reinterpret_cast<u4v*>(&Hh[pidx(row, kv * 8)]) = val;
        }
    }
    f4v acc2[MI][J];
    zero(acc2);
    bf8v v1[K4D][JF], w2[K4F][J];
    load_w<JF, K4D>(v1, a.W1, D, 0, 0, wf, lane);
    __syncthreads();
    for (int c0 = 0; c0 < ff; c0 += FC) {
        load_w<J, K4F>(w2, a.W2, ff, 0, c0, wn, lane);  // in flight during ff1
        f4v acc1[MI][JF];
        zero(acc1);
        mma_panel<SPLIT, MI, JF, K4D>(acc1, Hh, Hl, v1, r16, kq);
        const f4v bb = *reinterpret_cast<const f4v*>(a.b1 + c0 + wf + c4);
        if (c0 + FC < ff) load_w<JF, K4D>(v1, a.W1, D, c0 + FC, 0, wf, lane);  // in flight during GELU and ff2
        __syncthreads();  // every wave has finished reading the previous intermediate chunk
#pragma unroll
        for (int i = 0; i < MI; i++) {
            f4v v;
#pragma unroll
            for (int r = 0; r < 4; r++) v[r] = gelu_fast(acc1[i][0][r] + bb[r]);
            put_a4<SPLIT>(Fh, Fl, i * 16 + r16, wf + c4, v);
        }
        __syncthreads();
        mma_panel<SPLIT, MI, J, K4F>(acc2, Fh, Fl, w2, r16, kq);
    }
    const auto rX = rows_rsrc(a.X + (long)m0 * D, rows * D * 4);
#pragma unroll
    for (int j = 0; j < J; j++) {
        const f4v b = *reinterpret_cast<const f4v*>(a.b2 + wn + j * 16 + c4);
#pragma unroll
        for (int i = 0; i < MI; i++) {
            const int o = ((i * 16 + r16) * D + wn + j * 16 + c4) * 4;
            f4v x = bload4(rX, o);
#pragma unroll
            for (int r = 0; r < 4; r++) x[r] = x[r] + acc2[i][j][r] + b[r];
            bstore4(rX, o, x);
        }
    }
}

// One block of the wider trunk (D = 256) between two attentions, over
// 64-token panels, eight waves (32 columns of every D-column product and 16
// of every 128-column ff1 chunk each): X += O @ Wp + bp; H = LN2(X);
// X += ff2(GELU(ff1(H))) + b2 (X stored); with QKV, H = LN1'(X) and
// QKV = H @ Wqkv' + bqkv'. Every product is the tile GEMM's dot product in
// its order, with its bias and residual arithmetic, and each LayerNorm's
// statistics are taken as k_layernorm takes them (the panel's residual rows
// go through LDS so one wave reads a row as k_layernorm's lanes do), so the
// launch gives the same bits as the tile GEMMs, the MLP launch and the
// LayerNorm launches it replaces. LDS: one 64 KB panel (O, then X as f32
// for the statistics, then H) and one 128-column intermediate chunk (96 KB
// split: one block per CU).
template <bool SPLIT, int D>
__device__ __forceinline__ void ln_panel_k(const f4v (&xm)[4][D / 128], float* Xs, float* mrow, float* irow, int wn,
                                           int r16, int c4, int wave, int lane) {
#pragma clang fp contract(off)
    constexpr int J = D / 128;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < J; j++)
            *reinterpret_cast<f4v*>(&Xs[(i * 16 + r16) * D + wn + j * 16 + c4]) = xm[i][j];
    __syncthreads();
    // k_layernorm's order: lane l adds d = l + 64 k in k order, then the
    // butterfly; the rows of the panel go to the eight waves in turn
#pragma unroll 1
    for (int r = wave; r < RB; r += 8) {
        const float* xr = Xs + r * D;
        float v[D / 64];
        float sum = 0.f;
#pragma unroll
        for (int k = 0; k < D / 64; k++) {
            v[k] = xr[lane + 64 * k];
            sum += v[k];
        }
        const float mean = wave_sum(sum) / (float)D;
        float q = 0.f;
#pragma unroll
        for (int k = 0; k < D / 64; k++) {
            const float c = v[k] - mean;
            q += c * c;
        }
        const float inv = 1.0f / sqrtf(wave_sum(q) / (float)D + 1e-5f);
        if (lane == 0) {
            mrow[r] = mean;
            irow[r] = inv;
        }
    }
    __syncthreads();  // every read of Xs is done; the statistics are in
}

// QKV = H @ Wqkv + bqkv for a 64-token panel of d = 256 (H in the panel
// Ah / Al), eight waves of 32 columns, three 256-column chunks; w holds the
// first chunk's fragments on entry. f32 (split) or bf16 out.
template <bool SPLIT>
__device__ __forceinline__ void qkv_panel256(const uint16_t* Ah, const uint16_t* Al, bf8v (&w)[8][2],
                                             const RowsArgs& a, int m0, int rows, int wn, int r16, int c4, int kq,
                                             int lane) {
    constexpr int D = 256, MI = 4, J = 2, K4D = 8;
    const auto rQ = rows_rsrc(static_cast<const char*>(a.QKV) + (long)m0 * 3 * D * (SPLIT ? 4 : 2),
                              rows * 3 * D * (SPLIT ? 4 : 2));
#pragma unroll 1
    for (int c = 0; c < 3; c++) {
        asm volatile("" ::: "memory");  // H's fragments read per chunk, not held across all three
        f4v acc[MI][J];
        zero(acc);
        mma_panel<SPLIT, MI, J, K4D>(acc, Ah, Al, w, r16, kq);
        if (c < 2) load_w<J, K4D>(w, a.Wqkv, D, (c + 1) * D, 0, wn, lane);
#pragma unroll
        for (int j = 0; j < J; j++) {
            const f4v b = *reinterpret_cast<const f4v*>(a.bqkv + c * D + wn + j * 16 + c4);
#pragma unroll
            for (int i = 0; i < MI; i++) {
                f4v v;
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = acc[i][j][r] + b[r];
                const int o = (((i * 16 + r16) * 3 + c) * D + wn + j * 16 + c4) * 4;
                if (SPLIT) bstore4(rQ, o, v);
                else bstore4(rQ, o / 2, bf4(v));
            }
        }
    }
}

template <bool SPLIT, bool QKV>
__global__ __launch_bounds__(512, 1) void k_rows256(int M, int ff, RowsArgs a) {
    constexpr int D = 256, WN = 8, NT = 64 * WN, MI = 4, J = D / (16 * WN), P = SPLIT ? 2 : 1;
    constexpr int FC = 128, K4D = D / 32, K4F = FC / 32;
    static_assert(J == 2, "eight waves of 32 columns");
    constexpr int PANEL = RB * D * 4;  // bytes: the f32 residual rows, or the bf16 (pair) panel
    __shared__ __align__(16) unsigned char As[PANEL];
    __shared__ __align__(16) uint16_t Fs[P * RB * FC];
    __shared__ float mrow[RB], irow[RB];
    uint16_t* const Ah = reinterpret_cast<uint16_t*>(As);
    uint16_t* const Al = Ah + (SPLIT ? RB * D : 0);
    float* const Xs = reinterpret_cast<float*>(As);
    uint16_t *const Fh = Fs, *const Fl = Fs + (SPLIT ? RB * FC : 0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * RB, rows = min(RB, M - m0);
    const int wn = wave * 16 * J, wf = wave * 16;
    const int r16 = lane & 15, kq = (lane >> 4) * 8, c4 = (lane >> 4) * 4;
    const auto rX = rows_rsrc(a.X + (long)m0 * D, rows * D * 4);
    auto at = [&](int i, int j) { return ((i * 16 + r16) * D + wn + j * 16 + c4) * 4; };
    // a panel of D columns from global (f32 split into hi + lo, or bf16) into Ah / Al
    auto stage = [&](const void* src) {
        if (SPLIT) {
            const auto r = rows_rsrc(static_cast<const float*>(src) + (long)m0 * D, rows * D * 4);
#pragma unroll
            for (int q = 0; q < RB * D / 4 / NT; q++) {
                const int v = tid + NT * q, row = v / (D / 4), kv = v % (D / 4);
                put_a4<true>(Ah, Al, row, kv * 4, bload4(r, (row * D + kv * 4) * 4));
            }
        } else {
            const auto r = rows_rsrc(static_cast<const uint16_t*>(src) + (long)m0 * D, rows * D * 2);
#pragma unroll
            for (int q = 0; q < RB * D / 8 / NT; q++) {
                const int v = tid + NT * q, row = v / (D / 8), kv = v % (D / 8);
                const u4v val = __builtin_amdgcn_raw_buffer_load_b128(r, (row * D + kv * 8) * 2, 0, 0);
                *reinterpret_cast<u4v*>(&Ah[pidx(row, kv * 8)]) = val;
            }
        }
    };
    // H = LN(xm) with gamma g, beta b into Ah / Al (statistics in mrow / irow)
    auto put_ln = [&](const f4v(&xm)[MI][J], const float* g, const float* b) {
#pragma unroll
        for (int j = 0; j < J; j++) {
            const f4v gv = *reinterpret_cast<const f4v*>(g + wn + j * 16 + c4);
            const f4v bv = *reinterpret_cast<const f4v*>(b + wn + j * 16 + c4);
#pragma unroll
            for (int i = 0; i < MI; i++) {
                const int row = i * 16 + r16;
                const float mu = mrow[row], iv = irow[row];
                f4v h;
#pragma unroll
                for (int r = 0; r < 4; r++) h[r] = __builtin_fmaf((xm[i][j][r] - mu) * iv, gv[r], bv[r]);
                put_a4<SPLIT>(Ah, Al, row, wn + j * 16 + c4, h);
            }
        }
    };

    // ---- X += O @ Wp + bp
    bf8v w[K4D][J];
    stage(a.O);
    load_w<J, K4D>(w, a.Wp, D, 0, 0, wn, lane);
    __syncthreads();
    f4v xm[MI][J];
    {
        f4v acc[MI][J];
        zero(acc);
        mma_panel<SPLIT, MI, J, K4D>(acc, Ah, Al, w, r16, kq);
#pragma unroll
        for (int j = 0; j < J; j++) {
            const f4v b = *reinterpret_cast<const f4v*>(a.bp + wn + j * 16 + c4);
#pragma unroll
            for (int i = 0; i < MI; i++) {
                const f4v x = bload4(rX, at(i, j));
#pragma unroll
                for (int r = 0; r < 4; r++) xm[i][j][r] = x[r] + acc[i][j][r] + b[r];
            }
        }
    }
    __syncthreads();  // every read of O is done
    // ---- H = LN2(X)
    ln_panel_k<SPLIT, D>(xm, Xs, mrow, irow, wn, r16, c4, wave, lane);
    put_ln(xm, a.n2g, a.n2b);
    __syncthreads();

    // ---- X += ff2(GELU(ff1(H))) + b2, FC intermediate columns at a time
    f4v acc2[MI][J];
    zero(acc2);
    bf8v v1[K4D][1], w2[K4F][J];
    load_w<1, K4D>(v1, a.W1, D, 0, 0, wf, lane);
    for (int c0 = 0; c0 < ff; c0 += FC) {
        load_w<J, K4F>(w2, a.W2, ff, 0, c0, wn, lane);
        f4v acc1[MI][1];
        zero(acc1);
        mma_panel<SPLIT, MI, 1, K4D>(acc1, Ah, Al, v1, r16, kq);
        const f4v bb = *reinterpret_cast<const f4v*>(a.b1 + c0 + wf + c4);
        if (c0 + FC < ff) load_w<1, K4D>(v1, a.W1, D, c0 + FC, 0, wf, lane);
        __syncthreads();  // every wave has finished reading the previous intermediate chunk
#pragma unroll
        for (int i = 0; i < MI; i++) {
            f4v v;
#pragma unroll
            for (int r = 0; r < 4; r++) v[r] = gelu_fast(acc1[i][0][r] + bb[r]);
            put_a4<SPLIT>(Fh, Fl, i * 16 + r16, wf + c4, v);
        }
        __syncthreads();
        mma_panel<SPLIT, MI, J, K4F>(acc2, Fh, Fl, w2, r16, kq);
    }
#pragma unroll
    for (int j = 0; j < J; j++) {
        const f4v b = *reinterpret_cast<const f4v*>(a.b2 + wn + j * 16 + c4);
#pragma unroll
        for (int i = 0; i < MI; i++) {
#pragma unroll
            for (int r = 0; r < 4; r++) xm[i][j][r] = xm[i][j][r] + acc2[i][j][r] + b[r];
            bstore4(rX, at(i, j), xm[i][j]);
        }
    }
    if (!QKV) return;

    // ---- the next block's LN1 and its qkv product (every ff1 read of H is
    // done: the last chunk's barriers came after it)
    ln_panel_k<SPLIT, D>(xm, Xs, mrow, irow, wn, r16, c4, wave, lane);
    load_w<J, K4D>(w, a.Wqkv, D, 0, 0, wn, lane);
    put_ln(xm, a.n1g, a.n1b);
    __syncthreads();
    qkv_panel256<SPLIT>(Ah, Al, w, a, m0, rows, wn, r16, c4, kq, lane);
}

// The embedding, block 0's LN1 and its qkv product for d = 256, over
// 64-token panels of eight waves: each wave embeds two tokens at a time
// with 32 lanes a token (the arithmetic of k_cell_embed_ln_v), stores X and
// puts LN1's output straight into the panel; then the qkv product as in
// k_rows256. The same bits as the embedding launch and the qkv tile GEMM.
template <bool SPLIT>
__global__ __launch_bounds__(512, 1) void k_rows256_embed(int M, const float4* __restrict__ x4,
                                                          const float* __restrict__ we, const float* __restrict__ be,
                                                          const float* __restrict__ mvec,
                                                          const int* __restrict__ tok_req, RowsArgs a) {
    constexpr int D = 256, WN = 8, J = D / (16 * WN), K4D = D / 32;
    __shared__ __align__(16) unsigned char As[RB * D * 4];
    uint16_t* const Ah = reinterpret_cast<uint16_t*>(As);
    uint16_t* const Al = Ah + (SPLIT ? RB * D : 0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * RB, rows = min(RB, M - m0);
    const int wn = wave * 16 * J;
    const int r16 = lane & 15, kq = (lane >> 4) * 8, c4 = (lane >> 4) * 4;
    bf8v w[K4D][J];
    load_w<J, K4D>(w, a.Wqkv, D, 0, 0, wn, lane);
    const int j32 = lane & 31, d0 = j32 * 8;
#pragma unroll 1
    for (int pass = 0; pass < RB / 16; pass++) {
        const int row = pass * 16 + wave * 2 + (lane >> 5), m = m0 + row;
        float o[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (m < M) {  // both 32-lane halves of a token take the same branch
            float e[8];
            const float* mv = mvec + (tok_req ? (long)tok_req[m] * D : 0);
            embed_ln_token<D>(x4[m], we, be, mv, a.n1g, a.n1b, d0, e, o);
            float* xr = a.X + (long)m * D + d0;
            *reinterpret_cast<f4v*>(xr) = f4v{e[0], e[1], e[2], e[3]};
            *reinterpret_cast<f4v*>(xr + 4) = f4v{e[4], e[5], e[6], e[7]};
        }
        put_a4<SPLIT>(Ah, Al, row, d0, f4v{o[0], o[1], o[2], o[3]});
        put_a4<SPLIT>(Ah, Al, row, d0 + 4, f4v{o[4], o[5], o[6], o[7]});
    }
    __syncthreads();
    qkv_panel256<SPLIT>(Ah, Al, w, a, m0, rows, wn, r16, c4, kq, lane);
}

}  // namespace

bool rows_fused_ok(int dim, int ff) { return dim == RD && ff % RD == 0; }

void rows_fused(hipStream_t s, int M, int dim, int ff, bool split, const RowsArgs& a, int shape) {
    if (M <= 0) return;
    if (!rows_fused_ok(dim, ff)) fail(J2K_ERR_LIMIT, "fused row kernel needs dim == 128 and ff a multiple of 128");
    const bool proj = a.Wp != nullptr, qkv = a.Wqkv != nullptr, lnout = !qkv && a.n1g != nullptr;
    if (proj ? !a.O : !a.Hin) fail(J2K_ERR_INVALID, "fused row kernel: missing input");
    if (shape < ROWS_4W || shape > ROWS_2X2) fail(J2K_ERR_INVALID, "fused row kernel: unknown launch shape");
    const int waves = shape == ROWS_8W ? 8 : 4;
    dim3 grid(cdiv(M, RB)), block(64 * waves);
    const bool ro = a.ro.w != nullptr;
    if (ro && !(proj && !qkv && !lnout)) fail(J2K_ERR_INVALID, "fused row kernel: the readout ends the last block");
    if (ro) {
        if (split) hipLaunchKernelGGL((k_rows<true, true, false, false, 1, 4, 128, true>), grid, dim3(256), 0, s, M, ff, a);
        else hipLaunchKernelGGL((k_rows<false, true, false, false, 1, 4, 128, true>), grid, dim3(256), 0, s, M, ff, a);
        LAUNCHCHK("rows_fused");
        return;
    }
    if (a.emb.x4) {
        if (!(proj && qkv && shape == ROWS_4W) || !a.emb.we || !a.emb.be || !a.emb.mvec)
            fail(J2K_ERR_INVALID, "fused row kernel: the embedding enters the four-wave first block only");
        if (split)
            hipLaunchKernelGGL((k_rows<true, true, true, false, 1, 4, 128, false, true>), grid, dim3(256), 0, s, M, ff, a);
        else
            hipLaunchKernelGGL((k_rows<false, true, true, false, 1, 4, 128, false, true>), grid, dim3(256), 0, s, M, ff,
                               a);
        LAUNCHCHK("rows_fused");
        return;
    }
#define J2K_ROWS(S, P, Q, L)                                                                                          \
    if (split == S && proj == P && qkv == Q && lnout == L) {                                                          \
        if (shape == ROWS_4W) hipLaunchKernelGGL((k_rows<S, P, Q, L, 1, 4, 128>), grid, block, 0, s, M, ff, a);      \
        else if (shape == ROWS_8W) hipLaunchKernelGGL((k_rows<S, P, Q, L, 1, 8, 128>), grid, block, 0, s, M, ff, a); \
        else if (shape == ROWS_4W_F64) hipLaunchKernelGGL((k_rows<S, P, Q, L, 1, 4, 64>), grid, block, 0, s, M, ff, a); \
        else hipLaunchKernelGGL((k_rows<S, P, Q, L, 2, 2, 128>), grid, block, 0, s, M, ff, a);                        \
        LAUNCHCHK("rows_fused");                                                                                      \
        return;                                                                                                       \
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

void rows_embed_qkv(hipStream_t s, int M, int dim, bool split, const float* x4, const float* we, const float* be,
                    const float* mvec, const int* tok_req, const RowsArgs& a) {
    if (M <= 0) return;
    if (dim != RD) fail(J2K_ERR_LIMIT, "row-panel embedding kernel needs dim == 128");
    if (!a.n1g || !a.n1b || !a.Wqkv || !a.bqkv || !a.QKV || (a.n_zero > 0 && !a.zero))
        fail(J2K_ERR_INVALID, "row-panel embedding kernel: missing operand");
    if (split)
        hipLaunchKernelGGL(k_rows_embed_qkv<true>, dim3(cdiv(M, RB)), dim3(256), 0, s, M, (const float4*)x4, we, be,
                           mvec, tok_req, a);
    else
        hipLaunchKernelGGL(k_rows_embed_qkv<false>, dim3(cdiv(M, RB)), dim3(256), 0, s, M, (const float4*)x4, we, be,
                           mvec, tok_req, a);
    LAUNCHCHK("rows_embed_qkv");
}

void rows_readout(hipStream_t s, int M, int dim, const float* X, const Readout& ro) {
    if (M <= 0) return;
    if (dim != RD) fail(J2K_ERR_LIMIT, "row-panel readout needs dim == 128");
    if (!X || !ro.g || !ro.b || !ro.w || !ro.bias || !(ro.y || (ro.xn && ro.x0)))
        fail(J2K_ERR_INVALID, "row-panel readout: missing operand");
    hipLaunchKernelGGL(k_rows_readout, dim3(cdiv(M, RB)), dim3(256), 0, s, M, X, ro);
    LAUNCHCHK("rows_readout");
}

void rows_qkv(hipStream_t s, int M, int dim, bool split, const RowsArgs& a) {
    if (M <= 0) return;
    if (dim != RD) fail(J2K_ERR_LIMIT, "row-panel qkv kernel needs dim == 128");
    if (!a.Hin || !a.Wqkv || !a.bqkv || !a.QKV) fail(J2K_ERR_INVALID, "row-panel qkv kernel: missing operand");
    if (split) hipLaunchKernelGGL(k_rows_qkv<true>, dim3(cdiv(M, RB)), dim3(256), 0, s, M, a);
    else hipLaunchKernelGGL(k_rows_qkv<false>, dim3(cdiv(M, RB)), dim3(256), 0, s, M, a);
    LAUNCHCHK("rows_qkv");
}

bool mlp_fused_ok(int dim, int ff) { return dim == 256 && ff % 128 == 0; }

void rows256_embed_qkv(hipStream_t s, int M, int dim, bool split, const float* x4, const float* we, const float* be,
                       const float* mvec, const int* tok_req, const RowsArgs& a) {
    if (M <= 0) return;
    if (dim != 256) fail(J2K_ERR_LIMIT, "d = 256 embedding kernel needs dim == 256");
    if (!x4 || !we || !be || !mvec || !a.X || !a.n1g || !a.n1b || !a.Wqkv || !a.bqkv || !a.QKV)
        fail(J2K_ERR_INVALID, "d = 256 embedding kernel: missing operand");
    const dim3 grid(cdiv(M, RB)), block(512);
    if (split)
        hipLaunchKernelGGL(k_rows256_embed<true>, grid, block, 0, s, M, (const float4*)x4, we, be, mvec, tok_req, a);
    else
        hipLaunchKernelGGL(k_rows256_embed<false>, grid, block, 0, s, M, (const float4*)x4, we, be, mvec, tok_req, a);
    LAUNCHCHK("rows256_embed_qkv");
}

void rows256(hipStream_t s, int M, int dim, int ff, bool split, const RowsArgs& a) {
    if (M <= 0) return;
    if (!mlp_fused_ok(dim, ff)) fail(J2K_ERR_LIMIT, "d = 256 row-panel kernel needs dim == 256 and ff a multiple of 128");
    if (!a.O || !a.Wp || !a.bp || !a.n2g || !a.n2b || !a.W1 || !a.b1 || !a.W2 || !a.b2 || !a.X)
        fail(J2K_ERR_INVALID, "d = 256 row-panel kernel: missing operand");
    const bool qkv = a.Wqkv != nullptr;
    if (qkv && (!a.bqkv || !a.QKV || !a.n1g || !a.n1b)) fail(J2K_ERR_INVALID, "d = 256 row-panel kernel: missing qkv operand");
    const dim3 grid(cdiv(M, RB)), block(512);
    if (split) {
        if (qkv) hipLaunchKernelGGL((k_rows256<true, true>), grid, block, 0, s, M, ff, a);
        else hipLaunchKernelGGL((k_rows256<true, false>), grid, block, 0, s, M, ff, a);
    } else {
        if (qkv) hipLaunchKernelGGL((k_rows256<false, true>), grid, block, 0, s, M, ff, a);
        else hipLaunchKernelGGL((k_rows256<false, false>), grid, block, 0, s, M, ff, a);
    }
    LAUNCHCHK("rows256");
}

void mlp_fused(hipStream_t s, int M, int dim, int ff, bool split, const RowsArgs& a) {
    if (M <= 0) return;
    if (!mlp_fused_ok(dim, ff)) fail(J2K_ERR_LIMIT, "fused MLP kernel needs dim == 256 and ff a multiple of 128");
    if (!a.Hin || !a.W1 || !a.b1 || !a.W2 || !a.b2 || !a.X) fail(J2K_ERR_INVALID, "fused MLP kernel: missing operand");
    if (split) hipLaunchKernelGGL((k_mlp_rows<true, 256>), dim3(cdiv(M, RB)), dim3(512), 0, s, M, ff, a);
    else hipLaunchKernelGGL((k_mlp_rows<false, 256>), dim3(cdiv(M, RB)), dim3(512), 0, s, M, ff, a);
    LAUNCHCHK("mlp_fused");
}

}  // namespace j2k
