// Row statistics of a LayerNorm over 128 columns (cpu.rs:49-65: mean, then
// the biased variance of the centred values, 1/sqrt(var + 1e-5)) for rows
// whose values sit in MFMA accumulators: acc[i][j][r] is row wm + i*16 +
// 4*(lane/16) + r, column wn + j*16 + lane%16, spread over the block's waves.
//
// One fixed order for every kernel and block shape, so the tile GEMM
// epilogue and the row-panel kernel give bit-identical norms: each 16-column
// tile is summed as the butterfly xor 1, 2, 4, 8 over its columns sums it
// (((c0 + c1) + (c2 + c3)) + ...), then the 8 tile sums of a row are added in
// column order by one thread per row, which also takes the division and the
// square root once per row. Squares are rounded before they are summed (no
// contraction into fused multiply-adds), so every lane of a tile holds the
// same sum and either layout of the accumulators can reproduce it.
#pragma once

#include "common.h"

namespace j2k {

typedef float lnf4v __attribute__((ext_vector_type(4)));

// Sum over the 16 lanes of a column tile (the lanes sharing lane / 16), by
// the butterfly xor 1, 2, 4, 8: every lane of the tile ends with the total.
__device__ __forceinline__ float tile_sum(float x) {
#pragma clang fp contract(off)
    x += quad_xor1(x);
    x += quad_xor2(x);
    x += swz_xor<4>(x);
    x += swz_xor<8>(x);
    return x;
}

// mean and inv (1/sqrt(var + 1e-5)) of every row into mrow / irow [BM]
// (LDS, row = wm + i*16 + 4*(lane/16) + r); red is [8][BM] LDS. Synchronises
// the block four times; every thread of the block must call it. The
// statistics stay in LDS so they hold no registers across the caller.
template <int BM, int J>
__device__ __forceinline__ void ln_row_stats(const lnf4v (&acc)[4][J], int wm, int wn, int lane, float (*red)[BM],
                                             float* mrow, float* irow) {
#pragma clang fp contract(off)
    const int r16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = wm + i * 16 + g * 4 + r;
                const float mu = pass == 0 ? 0.f : mrow[row];
#pragma unroll
                for (int j = 0; j < J; j++) {
                    const float x = acc[i][j][r];
                    const float s = tile_sum(pass == 0 ? x : (x - mu) * (x - mu));
                    if (r16 == 0) red[wn / 16 + j][row] = s;
                }
            }
        __syncthreads();
        if ((int)threadIdx.x < BM) {
            float tot = 0.f;
#pragma unroll
            for (int t = 0; t < 8; t++) tot += red[t][threadIdx.x];
            if (pass == 0) mrow[threadIdx.x] = tot / 128.0f;
            else irow[threadIdx.x] = 1.0f / sqrtf(tot / 128.0f + 1e-5f);
        }
        __syncthreads();
    }
}

// The same statistics for rows held transposed: acc[i][j][r] (i < MI) is row
// wm + i*16 + lane%16, column wn + j*16 + 4*(lane/16) + r. A lane adds its
// four columns as the butterfly's first two levels do, (c0 + c1) + (c2 + c3);
// the four lanes of a row and tile leave their sums in red[tile][row][0..3]
// and the row's thread finishes the tree, (s0 + s1) + (s2 + s3), before the
// column-order sum over tiles. red is [8][BM][4] LDS (16-byte aligned).
template <int BM, int MI, int J>
__device__ __forceinline__ void ln_row_stats_t(const lnf4v (&acc)[MI][J], int wm, int wn, int lane,
                                               float (*red)[BM][4], float* mrow, float* irow) {
#pragma clang fp contract(off)
    const int r16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
#pragma unroll
        for (int i = 0; i < MI; i++) {
            const int row = wm + i * 16 + r16;
            const float mu = pass == 0 ? 0.f : mrow[row];
#pragma unroll
            for (int j = 0; j < J; j++) {
                lnf4v x = acc[i][j];
                if (pass == 1)
#pragma unroll
                    for (int r = 0; r < 4; r++) x[r] = (x[r] - mu) * (x[r] - mu);
                red[wn / 16 + j][row][g] = (x[0] + x[1]) + (x[2] + x[3]);
            }
        }
        __syncthreads();
        if ((int)threadIdx.x < BM) {
            float tot = 0.f;
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const lnf4v q = *reinterpret_cast<const lnf4v*>(red[t][threadIdx.x]);
                tot += (q[0] + q[1]) + (q[2] + q[3]);
            }
            if (pass == 0) mrow[threadIdx.x] = tot / 128.0f;
            else irow[threadIdx.x] = 1.0f / sqrtf(tot / 128.0f + 1e-5f);
        }
        __syncthreads();
    }
}

// Row totals over the 128 columns of values held as in ln_row_stats_t, in
// the same tree (the first pass of the statistics without the division):
// out[row] for every row of the block. Synchronises the block twice.
template <int BM, int MI, int J>
__device__ __forceinline__ void row_sums_t(const lnf4v (&val)[MI][J], int wm, int wn, int lane, float (*red)[BM][4],
                                           float* out) {
#pragma clang fp contract(off)
    const int r16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int i = 0; i < MI; i++)
#pragma unroll
        for (int j = 0; j < J; j++) {
            const lnf4v x = val[i][j];
            red[wn / 16 + j][wm + i * 16 + r16][g] = (x[0] + x[1]) + (x[2] + x[3]);
        }
    __syncthreads();
    if ((int)threadIdx.x < BM) {
        float tot = 0.f;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const lnf4v q = *reinterpret_cast<const lnf4v*>(red[t][threadIdx.x]);
            tot += (q[0] + q[1]) + (q[2] + q[3]);
        }
        out[threadIdx.x] = tot;
    }
    __syncthreads();
}

}  // namespace j2k
