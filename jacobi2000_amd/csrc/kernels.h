// Launchers for libj2k's HIP kernels. Every launcher enqueues on `s` and
// checks the launch; none synchronises.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace j2k {

enum Epi : int {
    EPI_STORE = 0,  // C = acc + bias
    EPI_GELU = 1,   // C = gelu(acc + bias)
    EPI_RESID = 2,  // C += acc + bias            (residual stream update)
    EPI_RESID_LN = 3,  // X += acc + bias, then H = LayerNorm(X row) in bf16
};

// ---- tiled bf16 GEMM (BM x 128 tiles, LDS-staged, XCD-aware order) ----------
// A [M, K] (bf16, or f32 for the split form), Bt [N, K] bf16, f32 bias.
// EPI_STORE/EPI_GELU write C [M, N]; EPI_RESID adds into the f32 stream X
// [M, N]; EPI_RESID_LN (N == 128) also writes H = LN(X) with gamma ln_g,
// beta ln_b. N % 16 == 0, K % 8 == 0. The split form feeds f32 activations
// to the bf16 MFMA as a hi + lo pair (two MFMAs per tile); its C, H are f32.
// `bm` is the block height, 64 or 128 (0: gemm_tile_default_bm).
bool gemm_tile_ok(int N, int K, bool split);
int gemm_tile_default_bm(int M, int N);
void gemm_bf16_tile(hipStream_t s, int M, int N, int K, const uint16_t* A, const uint16_t* Bt, const float* bias,
                    uint16_t* C, float* X, int epi, const float* ln_g, const float* ln_b, uint16_t* H, int bm = 0);
void gemm_bf16x2_tile(hipStream_t s, int M, int N, int K, const float* A, const uint16_t* Bt, const float* bias,
                      float* C, float* X, int epi, const float* ln_g, const float* ln_b, float* H, int bm = 0);

// FP8 form: A f32 quantized to E4M3 inside the kernel with each row's scale
// sa_m = rowmax[m] / 448 (rowmax from row_amax_f32 on the same stream), Bt
// E4M3 codes [N, K], C = (A_q @ Bt_q) * sa_m * cs[n] + bias with the same
// epilogues (C, H f32). N % 16 == 0, K % 16 == 0.
void gemm_f8_tile(hipStream_t s, int M, int N, int K, const float* A, const float* rowmax, const uint8_t* Bt,
                  const float* cs, const float* bias, float* C, float* X, int epi, const float* ln_g,
                  const float* ln_b, float* H, int bm = 0);
// out[m] = max |A[m, :]| over the K columns of each of the M rows of A.
void row_amax_f32(hipStream_t s, int M, int K, const float* A, float* out);

// ---- row-panel kernel (dim == 128, ff % 128 == 0) --------------------------
// Per 64-token panel: [X += O @ Wp + bp; H = LN2(X)] if Wp, else H = Hin;
// X += ff2(GELU(ff1(H))); then, if Wqkv, QKV = LN1'(X) @ Wqkv + bqkv, else
// if n1g, Hout = LN1'(X). Activations (O, Hin, Hout, QKV) are bf16, or f32
// with split; weights are the bf16 [N, K] copies.
// The cell_st readout of the served forms (final LayerNorm g, b, then
// y = LN(X) @ w [128][3] + bias): with xn set, xn = known ? x0 : x0 + y
// (float4 per token, known in .w); with y set, y [M][3].
struct Readout {
    const float *g = nullptr, *b = nullptr, *w = nullptr, *bias = nullptr;
    const float* x0 = nullptr;
    float* xn = nullptr;
    float* y = nullptr;
};
// Weights Wp, W1, W2, Wqkv: bf16 in MFMA fragment order (LayerW::*_f).
// The cell_st embedding's operands (X = embed(x4) + bias + meta vector of
// the token's request, k_cell_embed_ln's per-element arithmetic).
struct RowsEmbed {
    const float* x4 = nullptr;  // [M][4] token features
    const float *we = nullptr, *be = nullptr, *mvec = nullptr;
    const int* tok_req = nullptr;  // the token's request (its row of mvec); NULL: one request
};
struct RowsArgs {
    const void* O = nullptr;
    const uint16_t* Wp = nullptr;
    const float *bp = nullptr, *n2g = nullptr, *n2b = nullptr;
    const void* Hin = nullptr;
    const uint16_t *W1 = nullptr, *W2 = nullptr;
    const float *b1 = nullptr, *b2 = nullptr;
    float* X = nullptr;
    const float *n1g = nullptr, *n1b = nullptr;
    const uint16_t* Wqkv = nullptr;
    const float* bqkv = nullptr;
    void* QKV = nullptr;
    void* Hout = nullptr;
    Readout ro;  // last block with PROJ only: the readout instead of storing X
    // with emb.x4 (PROJ and QKV, four-wave shape): the residual entering the
    // block is the embedding, made here instead of read from X
    RowsEmbed emb;
    // the embedding kernel's first block also clears these words (the
    // residual and violation words a later projection accumulates into)
    unsigned* zero = nullptr;
    int n_zero = 0;
};
bool rows_fused_ok(int dim, int ff);
// Launch shapes of the row-panel kernel (same results in all): four waves of
// 32 product columns each with 128-column MLP chunks; eight waves of 16; four
// waves with 64-column MLP chunks (three blocks per CU); four waves as 2 x 2
// over 32 tokens and 64 columns each.
enum { ROWS_4W = 1, ROWS_8W = 2, ROWS_4W_F64 = 3, ROWS_2X2 = 4 };
void rows_fused(hipStream_t s, int M, int dim, int ff, bool split, const RowsArgs& a, int shape = ROWS_4W);
// X += ff2(GELU(ff1(Hin))) + biases for dim == 256 (Hin, W1, b1, W2, b2, X of
// a; weights in fragment order), the tile GEMMs' GELU and residual epilogue
// arithmetic with the intermediate kept on chip.
bool mlp_fused_ok(int dim, int ff);
void mlp_fused(hipStream_t s, int M, int dim, int ff, bool split, const RowsArgs& a);
// dim == 256, everything of a block between two attentions in one launch: X
// += O @ Wp + bp, H = LN2(X), X += ff2(GELU(ff1(H))) + b2 (stored), and with
// a.Wqkv the next block's LN1 and QKV; the same bits as the tile GEMMs, the
// fused MLP and k_layernorm (weights in fragment order).
void rows256(hipStream_t s, int M, int dim, int ff, bool split, const RowsArgs& a);
// dim == 256: the embedding (+ meta) and block 0's LN1 and qkv in one launch
// (a.X, a.n1g / n1b, a.Wqkv in fragment order, a.bqkv, a.QKV); the same bits
// as the embedding launch and the qkv tile GEMM.
void rows256_embed_qkv(hipStream_t s, int M, int dim, bool split, const float* x4, const float* we, const float* be,
                       const float* mvec, const int* tok_req, const RowsArgs& a);
// QKV = Hin @ Wqkv + bqkv alone (dim == 128), the row-panel kernel's
// arithmetic: the first layer's product after the embedding.
void rows_qkv(hipStream_t s, int M, int dim, bool split, const RowsArgs& a);
// The readout of X [M][128] alone, with the row-panel kernel's fused
// readout arithmetic (its order fixed by the LayerNorm statistics' tree).
void rows_readout(hipStream_t s, int M, int dim, const float* X, const Readout& ro);
// The cell_st embedding (x4 [M] float4, we [4][128], be, mvec per request
// through tok_req when not NULL) into X, the first block's LN1 (n1g, n1b)
// and QKV = LN1(X) @ Wqkv + bqkv, in one launch (dim == 128). With a.X
// NULL the embedding is not stored (the next row-panel launch remakes it
// through RowsArgs::emb).
void rows_embed_qkv(hipStream_t s, int M, int dim, bool split, const float* x4, const float* we, const float* be,
                    const float* mvec, const int* tok_req, const RowsArgs& a);

// ---- GEMM, f32 operands on v_mfma_f32_16x16x4_f32 ------------------------
// C[M,N] (ldc) = A[M,K] (lda) @ B[K,N] (row-major, ldb = N) with epilogue.
// A launch of few tiles splits K when ws holds gemm_f32_workspace(M, N, K)
// floats (same result on every run; without ws it runs in one pass).
void gemm_f32(hipStream_t s, int M, int N, int K, const float* A, int lda, const float* B, const float* bias, float* C,
              int ldc, int epi, float* ws = nullptr, size_t ws_floats = 0);
size_t gemm_f32_workspace(int M, int N, int K);
// Form P embed: X[b*tokens + t] = patches @ We + be + pos[t] + mvec.
void gemm_embed_p(hipStream_t s, int M, int N, int K, const float* A, const float* B, const float* bias,
                  const float* pos, int tokens, const float* mvec, float* C, float* ws = nullptr,
                  size_t ws_floats = 0);
// Zero-padded cross-correlation (cpu.rs:171-205) as an implicit GEMM:
// out[b, co, y, x] = bias[co] + sum w[co, ci, ky, kx] in[b, ci, y+ky-k/2, x+kx-k/2],
// in [B, cin, h, w] (channel stride `in_bstride` floats per batch item),
// optional GELU; with add_src, out = add_src + gelu?(conv).
// `in2`/`cin1`: channels >= cin1 come from in2 ([B, cin-cin1, h, w]) so a
// concatenated input (cpu.rs:340-342) needs no copy.
void conv_f32(hipStream_t s, int B, int cin, int h, int w, const float* in, const float* in2, int cin1,
              const float* wt, const float* bias, int cout, int k, float* out, int gelu, const float* add_src,
              float* ws = nullptr, size_t ws_floats = 0);

// ---- LayerNorm (cpu.rs:49-65): one wave per row; f32 and/or bf16 out. ----
void layernorm(hipStream_t s, int rows, int dim, const float* x, const float* g, const float* b, float* out,
               uint16_t* out_bf);

// ---- Stencil gather attention -------------------------------------------
// Token geometry of one segment (a batch item or a request): tokens
// [base, base + T*ny*nx), index t*ny*nx + y*nx + x.
struct Seg {
    int base, T, ny, nx;
};
enum AttnKind : int {
    ATTN_PATCH_STENCIL = 0,  // mask.rs:18-30 on the token grid, clipped, dilated
    ATTN_FULL = 1,           // every token of the segment
    ATTN_CELL_ST = 2,        // (2r+1)^2 periodic in space x (2rt+1) clipped in time
};
struct AttnGeo {
    int kind, radius, dilation, time_radius;
};
// A block covers `tokens_per_block` consecutive tokens of one segment: tiles[i] = {seg, first local token}.
struct Tile {
    int seg, start;
};
// qkv [M, 3*dim] (q | k | v per row, heads contiguous within each), out [M, dim].
void attention_f32(hipStream_t s, const float* qkv, float* out, int dim, int heads, AttnGeo geo, const Seg* segs,
                   const Tile* tiles, int n_tiles, int tokens_per_block);
void attention_bf16(hipStream_t s, const uint16_t* qkv, uint16_t* out, int dim, int heads, AttnGeo geo,
                    const Seg* segs, const Tile* tiles, int n_tiles, int tokens_per_block);
int attn_tokens_per_block(int dim);  // tokens of one block of the generic attention (all heads)

// cell_st attention with K/V staged in LDS: one block per 8 x 8 tile of cells
// of one time slice ({segment, t, y0, x0}), periodic halo `rs`, time window
// rt. Staging modes (same results): K and V together, K then V in half the
// LDS, or one time plane at a time.
struct CellTile {
    int seg, t, y0, x0;
};
// Staging modes of the tiled cell attention (the same result in every one).
enum KvMode : int { KV_JOINT = 0, KV_SPLIT = 1, KV_PLANE = 2, KV_PAIR2 = 3, KV_PAIR3 = 4, KV_MODES = 5 };
void build_cell_tiles(const std::vector<Seg>& segs, std::vector<CellTile>& tiles);
size_t attn_cell_smem(int hd, int rs, int rt, int elem, int mode);
bool attn_cell_ok(int hd, int rs, int rt, int elem, int mode);  // a tiled kernel exists and fits in LDS
void attn_cell_prepare();  // once per device, before the first launch
void attention_cell_f32(hipStream_t s, const float* qkv, float* out, int dim, int heads, int rs, int rt,
                        const Seg* segs, const CellTile* tiles, int n_tiles, int mode);
void attention_cell_bf16(hipStream_t s, const uint16_t* qkv, uint16_t* out, int dim, int heads, int rs, int rt,
                         const Seg* segs, const CellTile* tiles, int n_tiles, int mode);

// The attention tables of one packed batch; cell tiles are used when present.
struct AttnPlan {
    const Seg* segs = nullptr;
    const Tile* tiles = nullptr;
    int n_tiles = 0;
    const CellTile* ctiles = nullptr;
    int n_ctiles = 0;
};

// ---- patch2d helpers -----------------------------------------------------
// cpu.rs:253-269: patches[b*tokens + t, c*p*p + py*p + px].
void patchify(hipStream_t s, int B, int C, int ny, int nx, int p, const float* state, float* patches);
// cpu.rs:276-277: mvec = meta @ Wm + bm (one row).
void meta_linear(hipStream_t s, int meta_dim, int dim, const float* meta, const float* w, const float* b, float* out);
// meta_linear of n vectors meta [n][meta_dim] into out [n][dim], one launch.
void meta_linear_n(hipStream_t s, int n, int meta_dim, int dim, const float* meta, const float* w, const float* b,
                   float* out);
// cpu.rs:319-324: xn [B, tokens, dim] -> t0 [B, dim, tokens].
void transpose_td(hipStream_t s, int B, int tokens, int dim, const float* in, float* out);
// cpu.rs:211-245, both axes: in [B, c, h, w] -> out [B, c, h*p, w*p] (out batch stride given).
void upsample2(hipStream_t s, int B, int c, int h, int w, int p, const float* in, float* out, long out_bstride);

// ---- projections ---------------------------------------------------------
// Per grid b of a batch [B, 3, n]: project_mass to target (mass_t[b] or,
// when mass_t is null, mass of `ref`), project_energy against prev energy
// (eprev[b]), then store the energy after projection in eout[b] (f64),
// violation terms into viol[b]. Sums in f64. (project.rs:9-54, eval.rs:144-156)
void project_p(hipStream_t s, int B, int n, float* x, const double* mass_t, int do_mass, const double* eprev,
               int do_energy, double* eout, float g, const float* mean, const float* stdv, float depth,
               float r_max);
// f64 sums of mass (mean of channel 0) and energy per grid b.
void grid_sums(hipStream_t s, int B, int n, const float* x, double* mass, double* energy, float g, const float* mean,
               const float* stdv, float depth);
// x[i] += y[i]
void add_inplace(hipStream_t s, long n, float* x, const float* y);

// ---- cell_st -----------------------------------------------------------
// Per slice: {first token, cells, request}.
struct Slice {
    int start, n, req;
};
// [3,T,ny,nx] planes <-> token-major float4 (h, u, v, known); the pack also
// writes `copy` when it is not NULL.
void cell_pack(hipStream_t s, int M, const float* grid, const uint8_t* known, float* x4, float* copy = nullptr);
void cell_unpack(hipStream_t s, int M, const float* x4, float* grid);
// Many requests in one launch each (the server's turn): job j moves `tokens`
// tokens between byte offset `buf` of a staging buffer (the request's grid
// planes [3][tokens] f32, for the pack followed by its mask [tokens] u8) and
// packed token `tok` (x4 / copy as in cell_pack and cell_unpack).
struct CellJob {
    long buf;
    int tok, tokens;
};
void cell_pack_jobs(hipStream_t s, int n_jobs, const CellJob* jobs, int max_tokens, const unsigned char* staging,
                    float* x4, float* copy);
void cell_unpack_jobs(hipStream_t s, int n_jobs, const CellJob* jobs, int max_tokens, const float* x4,
                      unsigned char* staging);
// Repacking: job j moves tokens [buf, buf + tokens) of x0_old and x_old to
// [tok, tok + tokens) of x0 and x (buf is a token index here).
void cell_move_jobs(hipStream_t s, int n_jobs, const CellJob* jobs, int max_tokens, const float* x0_old,
                    const float* x_old, float* x0, float* x);
// The request index of every token of a packing: tok_req[t] = r for t in segment r.
void seg_tokens(hipStream_t s, int n_segs, const Seg* segs, int* tok_req);
// A chained rollout window of T slices of n cells from the packed anchor
// slice anchor4 [n] float4: every slice copies its (h, u, v), slice 0 known;
// written to x0 and x1 (the anchor must not overlap them).
void cell_window(hipStream_t s, int n, int T, const float* anchor4, float* x0, float* x1);
// X = embed([h,u,v,known]) + mvec (mvec per request), H = LN1_0(X).
void cell_embed_ln(hipStream_t s, int M, int dim, const float* x4, const float* we, const float* be,
                   const float* mvec, const int* tok_req, const float* g, const float* b, float* X, float* H,
                   uint16_t* Hb);
// y = LN(X) @ Wr + br; xn = known ? x0 : x0 + y (either output may be null).
void cell_readout(hipStream_t s, int M, int dim, const float* X, const float* g, const float* b, const float* wr,
                  const float* br, const float* x0, float* xn, float* y);
// One request (its slices in grid order): cell_pack into x0 and x1 and
// cell_targets of the result, in one launch (the same values); with meta,
// also meta_linear into mvec.
void cell_pack_targets(hipStream_t s, int n_slices, const Slice* sl, int M, const float* grid, const uint8_t* known,
                       float* x0, float* x1, double* mass_t, double* eref_t, float g, const float* mean,
                       const float* stdv, float depth, int meta_dim = 0, int dim = 0, const float* meta = nullptr,
                       const float* meta_w = nullptr, const float* meta_b = nullptr, float* mvec = nullptr);
// Per-slice targets from x0: mass (mean normalized h) and energy (f64).
void cell_targets(hipStream_t s, int n_slices, const Slice* sl, const float* x0, double* mass_t, double* eref_t,
                  float g, const float* mean, const float* stdv, float depth);
// Projection of every slice on its blank cells, then max|xn - xc| over blank
// cells into resid_bits[req] (float bits, atomic max; caller zeroes it) and,
// when viol_bits is not NULL, cell_violation of the result into it. With
// grid (a single request of M cells), the result also goes to grid [3][M]
// as cell_unpack writes it.
void cell_project(hipStream_t s, int n_slices, const Slice* sl, float* xn, const float* xc, const double* mass_t,
                  const double* eref_t, unsigned* resid_bits, float g, const float* mean, const float* stdv,
                  float depth, float r_max, unsigned long long* viol_bits = nullptr, float* grid = nullptr,
                  int M = 0);
// Per-slice relative energy violation, max into viol_bits[req] (double bits,
// atomic max). Slice i's reference energy is eref_t[i], or eref_t[eref_idx[i]]
// when eref_idx is given.
void cell_violation(hipStream_t s, int n_slices, const Slice* sl, const float* x4, const double* eref_t,
                    unsigned long long* viol_bits, float g, const float* mean, const float* stdv, float depth,
                    float r_max, const int* eref_idx = nullptr);

}  // namespace j2k

namespace j2k {
// p[i] = v for i < n (a kernel, so it can sit inside a captured graph).
void fill_u32(hipStream_t s, unsigned* p, int n, unsigned v);
}  // namespace j2k
