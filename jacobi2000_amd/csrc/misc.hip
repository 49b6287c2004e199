// LayerNorm, patch2d glue kernels (patchify, meta linear, transpose,
// upsample) and the cell_st per-token kernels (pack, embed+LN, readout).
#include "common.h"
#include "embedln.h"
#include "kernels.h"

namespace j2k {

// ---- LayerNorm: one wave per row, cpu.rs:49-65 (two-pass mean/variance,
// biased, eps 1e-5 inside the sqrt). dim <= 1024. Each lane adds its values
// d = lane + 64 i in order, the wave butterfly (wave_sum) adds the lanes;
// squares are rounded before they are added; the output is
// fma((v - mean) * inv, gamma, beta). The d = 256 row-panel kernel
// computes the same statistics in the same order.
__global__ __launch_bounds__(256) void k_layernorm(int rows, int dim, const float* __restrict__ x,
                                                   const float* __restrict__ g, const float* __restrict__ b,
                                                   float* __restrict__ out, uint16_t* __restrict__ out_bf) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rows) return;
    const float* xr = x + (long)row * dim;
    float v[16];
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int d = lane + i * 64;
        v[i] = d < dim ? xr[d] : 0.f;
        s += v[i];
    }
    const float mean = wave_sum(s) / (float)dim;
    float q = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int d = lane + i * 64;
        float c = v[i] - mean;
        q += d < dim ? c * c : 0.f;
    }
    const float inv = 1.0f / sqrtf(wave_sum(q) / (float)dim + 1e-5f);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int d = lane + i * 64;
        if (d < dim) {
            const float o = __builtin_fmaf((v[i] - mean) * inv, g[d], b[d]);
            if (out) out[(long)row * dim + d] = o;
            if (out_bf) out_bf[(long)row * dim + d] = f2bf(o);
        }
    }
}

void layernorm(hipStream_t s, int rows, int dim, const float* x, const float* g, const float* b, float* out,
               uint16_t* out_bf) {
    if (dim > 1024) fail(J2K_ERR_LIMIT, "layernorm: dim > 1024 is not supported");
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_layernorm, dim3(cdiv(rows, 4)), dim3(256), 0, s, rows, dim, x, g, b, out, out_bf);
    LAUNCHCHK("layernorm");
}

// ---- patch2d glue ----------------------------------------------------------

__global__ void k_patchify(int B, int C, int ny, int nx, int p, const float* __restrict__ st, float* __restrict__ pt) {
    const int tx = nx / p, ty = ny / p, tokens = tx * ty, pd = C * p * p;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x, total = (long)B * tokens * pd;
    if (i >= total) return;
    int d = (int)(i % pd);
    long bt = i / pd;
    int t = (int)(bt % tokens), b = (int)(bt / tokens);
    int c = d / (p * p), r = d % (p * p), py = r / p, px = r % p;
    int y = (t / tx) * p + py, x = (t % tx) * p + px;
    pt[i] = st[(((long)b * C + c) * ny + y) * nx + x];
}

void patchify(hipStream_t s, int B, int C, int ny, int nx, int p, const float* state, float* patches) {
    long total = (long)B * (nx / p) * (ny / p) * C * p * p;
    hipLaunchKernelGGL(k_patchify, dim3(cdiv(total, 256)), dim3(256), 0, s, B, C, ny, nx, p, state, patches);
    LAUNCHCHK("patchify");
}

// cpu.rs:26-45 for one row: bias first, then inputs in order.
__global__ void k_meta_linear(int md, int dim, const float* meta, const float* w, const float* b, float* out) {
    int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= dim) return;
    out[o] = meta_out(md, dim, meta, w, b, o);
}

void meta_linear(hipStream_t s, int meta_dim, int dim, const float* meta, const float* w, const float* b, float* out) {
    hipLaunchKernelGGL(k_meta_linear, dim3(cdiv(dim, 256)), dim3(256), 0, s, meta_dim, dim, meta, w, b, out);
    LAUNCHCHK("meta_linear");
}

__global__ void k_meta_linear_n(int md, int dim, const float* meta, const float* w, const float* b, float* out) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= dim) return;
    out[(long)blockIdx.y * dim + o] = meta_out(md, dim, meta + (long)blockIdx.y * md, w, b, o);
}

void meta_linear_n(hipStream_t s, int n, int meta_dim, int dim, const float* meta, const float* w, const float* b,
                   float* out) {
    for (int i0 = 0; i0 < n; i0 += 65535) {
        const int k = std::min(65535, n - i0);
        hipLaunchKernelGGL(k_meta_linear_n, dim3(cdiv(dim, 256), k), dim3(256), 0, s, meta_dim, dim,
                           meta + (long)i0 * meta_dim, w, b, out + (long)i0 * dim);
        LAUNCHCHK("meta_linear_n");
    }
}

__global__ void k_transpose_td(int B, int tokens, int dim, const float* __restrict__ in, float* __restrict__ out) {
    __shared__ float tile[32][33];
    const int b = blockIdx.z, t0 = blockIdx.x * 32, d0 = blockIdx.y * 32;
    const float* ib = in + (long)b * tokens * dim;
    float* ob = out + (long)b * tokens * dim;
    for (int r = threadIdx.y; r < 32; r += 8) {
        int t = t0 + r, d = d0 + threadIdx.x;
        if (t < tokens && d < dim) tile[r][threadIdx.x] = ib[(long)t * dim + d];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {
        int d = d0 + r, t = t0 + threadIdx.x;
        if (t < tokens && d < dim) ob[(long)d * tokens + t] = tile[threadIdx.x][r];
    }
}

void transpose_td(hipStream_t s, int B, int tokens, int dim, const float* in, float* out) {
    dim3 grid(cdiv(tokens, 32), cdiv(dim, 32), B);
    hipLaunchKernelGGL(k_transpose_td, grid, dim3(32, 8), 0, s, B, tokens, dim, in, out);
    LAUNCHCHK("transpose_td");
}

// cpu.rs:211-245: upsample along y then along x, fused. The x pass reads the
// y-upsampled values a(1-w)+b*w of columns xt and min(xt+1, w-1).
__global__ void k_upsample2(int B, int c, int h, int w, int p, const float* __restrict__ in, float* __restrict__ out,
                            long out_bstride) {
    const int H = h * p, W = w * p;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x, total = (long)B * c * H * W;
    if (i >= total) return;
    int X = (int)(i % W);
    long r = i / W;
    int Y = (int)(r % H);
    long bc = r / H;
    int ci = (int)(bc % c), b = (int)(bc / c);
    const float* src = in + ((long)b * c + ci) * h * w;
    int yt = Y / p, iy = Y % p, ynext = yt + 1 < h - 1 ? yt + 1 : h - 1;
    int xt = X / p, ix = X % p, xnext = xt + 1 < w - 1 ? xt + 1 : w - 1;
    float wy = ((float)iy + 0.5f) / (float)p, wx = ((float)ix + 0.5f) / (float)p;
    float a = src[yt * w + xt] * (1.0f - wy) + src[ynext * w + xt] * wy;
    float bb = src[yt * w + xnext] * (1.0f - wy) + src[ynext * w + xnext] * wy;
    out[(long)b * out_bstride + ((long)ci * H + Y) * W + X] = a * (1.0f - wx) + bb * wx;
}

void upsample2(hipStream_t s, int B, int c, int h, int w, int p, const float* in, float* out, long out_bstride) {
    long total = (long)B * c * h * p * w * p;
    hipLaunchKernelGGL(k_upsample2, dim3(cdiv(total, 256)), dim3(256), 0, s, B, c, h, w, p, in, out, out_bstride);
    LAUNCHCHK("upsample2");
}

__global__ void k_add(long n, float* x, const float* y) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += y[i];
}

void add_inplace(hipStream_t s, long n, float* x, const float* y) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_add, dim3(cdiv(n, 256)), dim3(256), 0, s, n, x, y);
    LAUNCHCHK("add_inplace");
}

// ---- cell_st per-token kernels ---------------------------------------------

__global__ void k_cell_pack(int M, const float* __restrict__ g, const uint8_t* __restrict__ kn, float4* __restrict__ x4,
                            float4* __restrict__ copy) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const float4 v = make_float4(g[m], g[(long)M + m], g[2 * (long)M + m], kn[m] ? 1.0f : 0.0f);
    x4[m] = v;
    if (copy) copy[m] = v;
}

void cell_pack(hipStream_t s, int M, const float* grid, const uint8_t* known, float* x4, float* copy) {
    hipLaunchKernelGGL(k_cell_pack, dim3(cdiv(M, 256)), dim3(256), 0, s, M, grid, known, (float4*)x4, (float4*)copy);
    LAUNCHCHK("cell_pack");
}

__global__ void k_cell_unpack(int M, const float4* __restrict__ x4, float* __restrict__ g) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    float4 v = x4[m];
    g[m] = v.x;
    g[(long)M + m] = v.y;
    g[2 * (long)M + m] = v.z;
}

__global__ void k_cell_pack_jobs(const CellJob* __restrict__ jobs, const unsigned char* __restrict__ staging,
                                 float4* __restrict__ x4, float4* __restrict__ copy) {
    const CellJob j = jobs[blockIdx.y];
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= j.tokens) return;
    const float* g = reinterpret_cast<const float*>(staging + j.buf);
    const uint8_t* kn = staging + j.buf + 12L * j.tokens;
    const long M = j.tokens;
    const float4 v = make_float4(g[m], g[M + m], g[2 * M + m], kn[m] ? 1.0f : 0.0f);
    x4[(long)j.tok + m] = v;
    if (copy) copy[(long)j.tok + m] = v;
}

// job tables of more than 65535 entries go in several launches (grid y limit)
constexpr int JOBS_PER_LAUNCH = 65535;

void cell_pack_jobs(hipStream_t s, int n_jobs, const CellJob* jobs, int max_tokens, const unsigned char* staging,
                    float* x4, float* copy) {
    if (max_tokens <= 0) return;
    for (int j0 = 0; j0 < n_jobs; j0 += JOBS_PER_LAUNCH) {
        const int n = std::min(JOBS_PER_LAUNCH, n_jobs - j0);
        hipLaunchKernelGGL(k_cell_pack_jobs, dim3(cdiv(max_tokens, 256), n), dim3(256), 0, s, jobs + j0, staging,
                           (float4*)x4, (float4*)copy);
        LAUNCHCHK("cell_pack_jobs");
    }
}

// job j: tokens [buf, buf + tokens) of x0_old and x_old to [tok, tok + tokens) of x0 and x
__global__ void k_cell_move_jobs(const CellJob* __restrict__ jobs, const float4* __restrict__ x0_old,
                                 const float4* __restrict__ x_old, float4* __restrict__ x0, float4* __restrict__ x) {
    const CellJob j = jobs[blockIdx.y];
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= j.tokens) return;
    x0[(long)j.tok + m] = x0_old[j.buf + m];
    x[(long)j.tok + m] = x_old[j.buf + m];
}

void cell_move_jobs(hipStream_t s, int n_jobs, const CellJob* jobs, int max_tokens, const float* x0_old,
                    const float* x_old, float* x0, float* x) {
    if (max_tokens <= 0) return;
    for (int j0 = 0; j0 < n_jobs; j0 += JOBS_PER_LAUNCH) {
        const int n = std::min(JOBS_PER_LAUNCH, n_jobs - j0);
        hipLaunchKernelGGL(k_cell_move_jobs, dim3(cdiv(max_tokens, 256), n), dim3(256), 0, s, jobs + j0,
                           (const float4*)x0_old, (const float4*)x_old, (float4*)x0, (float4*)x);
        LAUNCHCHK("cell_move_jobs");
    }
}

__global__ void k_cell_unpack_jobs(const CellJob* __restrict__ jobs, const float4* __restrict__ x4,
                                   unsigned char* __restrict__ staging) {
    const CellJob j = jobs[blockIdx.y];
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= j.tokens) return;
    float* g = reinterpret_cast<float*>(staging + j.buf);
    const long M = j.tokens;
    const float4 v = x4[(long)j.tok + m];
    g[m] = v.x;
    g[M + m] = v.y;
    g[2 * M + m] = v.z;
}

void cell_unpack_jobs(hipStream_t s, int n_jobs, const CellJob* jobs, int max_tokens, const float* x4,
                      unsigned char* staging) {
    if (max_tokens <= 0) return;
    for (int j0 = 0; j0 < n_jobs; j0 += JOBS_PER_LAUNCH) {
        const int n = std::min(JOBS_PER_LAUNCH, n_jobs - j0);
        hipLaunchKernelGGL(k_cell_unpack_jobs, dim3(cdiv(max_tokens, 256), n), dim3(256), 0, s, jobs + j0,
                           (const float4*)x4, staging);
        LAUNCHCHK("cell_unpack_jobs");
    }
}

// A chained rollout's next window on the device: every slice of M = T * n
// tokens starts as the anchor slice's (h, u, v), slice 0 known, the others
// blank (the values a forward request's blank fill gives from that slice).
__global__ void k_cell_window(int n, int M, const float4* __restrict__ anchor, float4* __restrict__ x0,
                              float4* __restrict__ x1) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    float4 v = anchor[m % n];
    v.w = m < n ? 1.0f : 0.0f;
    x0[m] = v;
    x1[m] = v;
}

void cell_window(hipStream_t s, int n, int T, const float* anchor4, float* x0, float* x1) {
    const int M = n * T;
    hipLaunchKernelGGL(k_cell_window, dim3(cdiv(M, 256)), dim3(256), 0, s, n, M, (const float4*)anchor4,
                       (float4*)x0, (float4*)x1);
    LAUNCHCHK("cell_window");
}

// tok_req[t] = r for every token t of segment r: one block per segment.
__global__ void k_seg_tokens(const Seg* __restrict__ segs, int* __restrict__ tok_req) {
    const Seg g = segs[blockIdx.x];
    const int n = g.T * g.ny * g.nx;
    for (int i = threadIdx.x; i < n; i += blockDim.x) tok_req[g.base + i] = blockIdx.x;
}

void seg_tokens(hipStream_t s, int n_segs, const Seg* segs, int* tok_req) {
    if (n_segs <= 0) return;
    hipLaunchKernelGGL(k_seg_tokens, dim3(n_segs), dim3(1024), 0, s, segs, tok_req);
    LAUNCHCHK("seg_tokens");
}

void cell_unpack(hipStream_t s, int M, const float* x4, float* grid) {
    hipLaunchKernelGGL(k_cell_unpack, dim3(cdiv(M, 256)), dim3(256), 0, s, M, (const float4*)x4, grid);
    LAUNCHCHK("cell_unpack");
}

// One wave per token: X = be + sum_i f_i We[i] (inputs in order, as cpu.rs:26-45)
// + mvec, then H = LN(X) with the first block's norm1.
__global__ __launch_bounds__(256) void k_cell_embed_ln(int M, int dim, const float4* __restrict__ x4,
                                                       const float* __restrict__ we, const float* __restrict__ be,
                                                       const float* __restrict__ mvec, const int* __restrict__ tok_req,
                                                       const float* __restrict__ g, const float* __restrict__ b,
                                                       float* __restrict__ X, float* __restrict__ H,
                                                       uint16_t* __restrict__ Hb) {
    const int lane = threadIdx.x & 63;
    const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (m >= M) return;
    const float4 f = x4[m];
    const float* mv = mvec + (tok_req ? (long)tok_req[m] * dim : 0);
    float v[16];
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int d = lane + i * 64;
        float a = 0.f;
        if (d < dim) {
            a = be[d];
            if (f.x != 0.f) a += f.x * we[d];
            if (f.y != 0.f) a += f.y * we[dim + d];
            if (f.z != 0.f) a += f.z * we[2 * dim + d];
            if (f.w != 0.f) a += f.w * we[3 * dim + d];
            a += mv[d];
            X[(long)m * dim + d] = a;
        }
        v[i] = a;
        s += a;
    }
    const float mean = wave_sum(s) / (float)dim;
    float q = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int d = lane + i * 64;
        float c = v[i] - mean;
        q += d < dim ? c * c : 0.f;
    }
    const float inv = 1.0f / sqrtf(wave_sum(q) / (float)dim + 1e-5f);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int d = lane + i * 64;
        if (d < dim) {
            float o = (v[i] - mean) * inv * g[d] + b[d];
            if (H) H[(long)m * dim + d] = o;
            if (Hb) Hb[(long)m * dim + d] = f2bf(o);
        }
    }
}

typedef float f4v __attribute__((ext_vector_type(4)));

// k_cell_embed_ln for dim = D (128 or 256): L = D/8 lanes per token, each
// with 8 contiguous columns (vector loads and stores), 64/L tokens per wave.
// Same per-element arithmetic as the generic kernel; the row sums group the
// lanes' partial sums.
template <int D>
__global__ __launch_bounds__(256) void k_cell_embed_ln_v(int M, const float4* __restrict__ x4,
                                                         const float* __restrict__ we, const float* __restrict__ be,
                                                         const float* __restrict__ mvec, const int* __restrict__ tok_req,
                                                         const float* __restrict__ g, const float* __restrict__ b,
                                                         float* __restrict__ X, float* __restrict__ H,
                                                         uint16_t* __restrict__ Hb) {
    constexpr int L = D / 8, TPW = 64 / L;
    const int lane = threadIdx.x & 63, j = lane % L;
    const int m = (blockIdx.x * 4 + (threadIdx.x >> 6)) * TPW + lane / L;
    if (m >= M) return;  // a whole group leaves together
    const float4 f = x4[m];
    const float* mv = mvec + (tok_req ? (long)tok_req[m] * D : 0);
    const int d0 = j * 8;
    float a[8], o[8];
    embed_ln_token<D>(f, we, be, mv, g, b, d0, a, o);
    float* xr = X + (long)m * D + d0;
    *reinterpret_cast<f4v*>(xr) = f4v{a[0], a[1], a[2], a[3]};
    *reinterpret_cast<f4v*>(xr + 4) = f4v{a[4], a[5], a[6], a[7]};
    if (H) {
        float* hr = H + (long)m * D + d0;
        *reinterpret_cast<f4v*>(hr) = f4v{o[0], o[1], o[2], o[3]};
        *reinterpret_cast<f4v*>(hr + 4) = f4v{o[4], o[5], o[6], o[7]};
    }
    if (Hb) {
        typedef uint32_t u4v __attribute__((ext_vector_type(4)));
        *reinterpret_cast<u4v*>(Hb + (long)m * D + d0) =
            u4v{f2bf2(o[0], o[1]), f2bf2(o[2], o[3]), f2bf2(o[4], o[5]), f2bf2(o[6], o[7])};
    }
}

void cell_embed_ln(hipStream_t s, int M, int dim, const float* x4, const float* we, const float* be, const float* mvec,
                   const int* tok_req, const float* g, const float* b, float* X, float* H, uint16_t* Hb) {
    if (dim > 1024) fail(J2K_ERR_LIMIT, "cell_embed_ln: dim > 1024 is not supported");
    if (dim == 128) {
        hipLaunchKernelGGL(k_cell_embed_ln_v<128>, dim3(cdiv(M, 16)), dim3(256), 0, s, M, (const float4*)x4, we, be,
                           mvec, tok_req, g, b, X, H, Hb);
    } else if (dim == 256) {
        hipLaunchKernelGGL(k_cell_embed_ln_v<256>, dim3(cdiv(M, 8)), dim3(256), 0, s, M, (const float4*)x4, we, be,
                           mvec, tok_req, g, b, X, H, Hb);
    } else {
        hipLaunchKernelGGL(k_cell_embed_ln, dim3(cdiv(M, 4)), dim3(256), 0, s, M, dim, (const float4*)x4, we, be, mvec,
                           tok_req, g, b, X, H, Hb);
    }
    LAUNCHCHK("cell_embed_ln");
}

// One wave per token: final LN of X, y = LN(X) @ Wr + br (3 outputs, wave
// reductions), xn = known ? x0 : x0 + y (known flag carried in .w).
__global__ __launch_bounds__(256) void k_cell_readout(int M, int dim, const float* __restrict__ X,
                                                      const float* __restrict__ g, const float* __restrict__ b,
                                                      const float* __restrict__ wr, const float* __restrict__ br,
                                                      const float4* __restrict__ x0, float4* __restrict__ xn,
                                                      float* __restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (m >= M) return;
    const float* xr = X + (long)m * dim;
    float v[16];
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int d = lane + i * 64;
        v[i] = d < dim ? xr[d] : 0.f;
        s += v[i];
    }
    const float mean = wave_sum(s) / (float)dim;
    float q = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int d = lane + i * 64;
        float c = v[i] - mean;
        q += d < dim ? c * c : 0.f;
    }
    const float inv = 1.0f / sqrtf(wave_sum(q) / (float)dim + 1e-5f);
    float y0 = 0.f, y1 = 0.f, y2 = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int d = lane + i * 64;
        if (d < dim) {
            float o = (v[i] - mean) * inv * g[d] + b[d];
            y0 += o * wr[d * 3 + 0];
            y1 += o * wr[d * 3 + 1];
            y2 += o * wr[d * 3 + 2];
        }
    }
    y0 = wave_sum(y0) + br[0];
    y1 = wave_sum(y1) + br[1];
    y2 = wave_sum(y2) + br[2];
    if (lane == 0) {
        if (y) {
            y[(long)m * 3] = y0;
            y[(long)m * 3 + 1] = y1;
            y[(long)m * 3 + 2] = y2;
        }
        if (xn) {
            float4 a = x0[m];
            xn[m] = a.w != 0.f ? a : make_float4(a.x + y0, a.y + y1, a.z + y2, 0.f);
        }
    }
}

// k_cell_readout for dim = D (128 or 256), L = D/8 lanes per token as in
// k_cell_embed_ln_v.
template <int D>
__global__ __launch_bounds__(256) void k_cell_readout_v(int M, const float* __restrict__ X,
                                                        const float* __restrict__ g, const float* __restrict__ b,
                                                        const float* __restrict__ wr, const float* __restrict__ br,
                                                        const float4* __restrict__ x0, float4* __restrict__ xn,
                                                        float* __restrict__ y) {
    constexpr int L = D / 8, TPW = 64 / L;
    const int lane = threadIdx.x & 63, j = lane % L;
    const int m = (blockIdx.x * 4 + (threadIdx.x >> 6)) * TPW + lane / L;
    if (m >= M) return;
    const int d0 = j * 8;
    const float* xr = X + (long)m * D + d0;
    const f4v lo = *reinterpret_cast<const f4v*>(xr), hi = *reinterpret_cast<const f4v*>(xr + 4);
    const float v[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) s += v[k];
    const float mean = group_sum<L>(s) / (float)D;
    float q = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float c = v[k] - mean;
        q += c * c;
    }
    const float inv = 1.0f / sqrtf(group_sum<L>(q) / (float)D + 1e-5f);
    float y0 = 0.f, y1 = 0.f, y2 = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int d = d0 + k;
        const float o = (v[k] - mean) * inv * g[d] + b[d];
        y0 += o * wr[d * 3 + 0];
        y1 += o * wr[d * 3 + 1];
        y2 += o * wr[d * 3 + 2];
    }
    y0 = group_sum<L>(y0) + br[0];
    y1 = group_sum<L>(y1) + br[1];
    y2 = group_sum<L>(y2) + br[2];
    if (j == 0) {
        if (y) {
            y[(long)m * 3] = y0;
            y[(long)m * 3 + 1] = y1;
            y[(long)m * 3 + 2] = y2;
        }
        if (xn) {
            const float4 a = x0[m];
            xn[m] = a.w != 0.f ? a : make_float4(a.x + y0, a.y + y1, a.z + y2, 0.f);
        }
    }
}

void cell_readout(hipStream_t s, int M, int dim, const float* X, const float* g, const float* b, const float* wr,
                  const float* br, const float* x0, float* xn, float* y) {
    if (dim > 1024) fail(J2K_ERR_LIMIT, "cell_readout: dim > 1024 is not supported");
    if (dim == 128) {
        hipLaunchKernelGGL(k_cell_readout_v<128>, dim3(cdiv(M, 16)), dim3(256), 0, s, M, X, g, b, wr, br,
                           (const float4*)x0, (float4*)xn, y);
    } else if (dim == 256) {
        hipLaunchKernelGGL(k_cell_readout_v<256>, dim3(cdiv(M, 8)), dim3(256), 0, s, M, X, g, b, wr, br,
                           (const float4*)x0, (float4*)xn, y);
    } else {
        hipLaunchKernelGGL(k_cell_readout, dim3(cdiv(M, 4)), dim3(256), 0, s, M, dim, X, g, b, wr, br,
                           (const float4*)x0, (float4*)xn, y);
    }
    LAUNCHCHK("cell_readout");
}

}  // namespace j2k

namespace j2k {

__global__ void k_fill_u32(unsigned* p, int n, unsigned v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void fill_u32(hipStream_t s, unsigned* p, int n, unsigned v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n, v);
    LAUNCHCHK("fill_u32");
}

}  // namespace j2k
