// Stencil attention of the cell_st form with K/V staged in LDS.
//
// A block owns an 8 x 8 tile of cells in one time slice of one segment and
// loops over the heads. For each head it stages the K and V rows of the
// tile's neighbourhood (the tile plus a periodic halo of `radius` cells in y
// and x, over the clipped time window t-rt .. t+rt) into LDS once, so every
// key row is read from L2 once per tile instead of once per query. Four lanes
// share a query, each holding a quarter of the head's dimensions: the dot
// products are reduced with two lane shuffles, the softmax runs single-pass
// with a running max in f32, and keys are visited in the order time offset,
// y offset, x offset (the order of the generic kernel and the oracle).
#include "common.h"
#include "kernels.h"

namespace j2k {

constexpr int CT_Y = 8, CT_X = 8;

template <class T>
__device__ __forceinline__ float ldf(const T* p);
template <>
__device__ __forceinline__ float ldf<float>(const float* p) {
    return *p;
}
template <>
__device__ __forceinline__ float ldf<uint16_t>(const uint16_t* p) {
    return bf2f(*p);
}

// n consecutive elements from LDS as f32, in 16-byte vectors.
template <int N>
__device__ __forceinline__ void lds_row(const float* p, float* o) {
#pragma unroll
    for (int i = 0; i < N; i += 4) {
        const float4 v = *reinterpret_cast<const float4*>(p + i);
        o[i] = v.x;
        o[i + 1] = v.y;
        o[i + 2] = v.z;
        o[i + 3] = v.w;
    }
}
template <int N>
__device__ __forceinline__ void lds_row(const uint16_t* p, float* o) {
    if constexpr (N % 8 == 0) {
#pragma unroll
        for (int i = 0; i < N; i += 8) {
            const uint4 v = *reinterpret_cast<const uint4*>(p + i);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                o[i + 2 * j] = __uint_as_float(w[j] << 16);
                o[i + 2 * j + 1] = __uint_as_float(w[j] & 0xffff0000u);
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) o[i] = bf2f(p[i]);
    }
}

template <int HD, class T>
__global__ __launch_bounds__(256) void k_attn_cell(const T* __restrict__ qkv, T* __restrict__ out, int dim, int heads,
                                                   int rs, int rt, const Seg* __restrict__ segs,
                                                   const CellTile* __restrict__ tiles) {
    constexpr int D4 = HD / 4;      // dims per lane
    constexpr int VEC = 16 / sizeof(T);  // elements per 16-byte vector
    extern __shared__ __align__(16) unsigned char smem[];
    const CellTile tl = tiles[blockIdx.x];
    const Seg sg = segs[tl.seg];
    const int hw = CT_X + 2 * rs, hh = CT_Y + 2 * rs;
    const int t_lo = max(tl.t - rt, 0), t_hi = min(tl.t + rt, sg.T - 1), nt = t_hi - t_lo + 1;
    const int ncell = nt * hh * hw;
    T* Ks = reinterpret_cast<T*>(smem);
    T* Vs = Ks + (size_t)ncell * HD;
    const long ld = 3L * dim;
    const int n = sg.ny * sg.nx;

    const int tid = threadIdx.x, q = tid >> 2, sub = tid & 3;
    const int qy = tl.y0 + (q >> 3), qx = tl.x0 + (q & 7);
    const bool live = qy < sg.ny && qx < sg.nx;
    const long qrow = (long)sg.base + (long)tl.t * n + (long)(live ? qy : 0) * sg.nx + (live ? qx : 0);
    const float scale = 1.0f / sqrtf((float)HD);

    for (int h = 0; h < heads; h++) {
        __syncthreads();  // the previous head's K/V are no longer read
        // stage K and V of head h for every cell of the neighbourhood
        constexpr int VPR = HD / VEC;  // vectors per row
        for (int v = tid; v < ncell * VPR * 2; v += 256) {
            const int which = v / (ncell * VPR), r = v - which * ncell * VPR;
            const int c = r / VPR, e = r - c * VPR;
            const int ct = c / (hh * hw), cr = c - ct * hh * hw, cy = cr / hw, cx = cr - cy * hw;
            int y = tl.y0 - rs + cy, x = tl.x0 - rs + cx;
            y = ((y % sg.ny) + sg.ny) % sg.ny;
            x = ((x % sg.nx) + sg.nx) % sg.nx;
            const long row = (long)sg.base + (long)(t_lo + ct) * n + (long)y * sg.nx + x;
            const uint4 val = *reinterpret_cast<const uint4*>(qkv + row * ld + (1 + which) * dim + h * HD + e * VEC);
            *reinterpret_cast<uint4*>((which ? Vs : Ks) + (size_t)c * HD + e * VEC) = val;
        }
        __syncthreads();
        float qv[D4], acc[D4], kv[D4];
#pragma unroll
        for (int d = 0; d < D4; d++) {
            qv[d] = ldf<T>(qkv + qrow * ld + h * HD + sub * D4 + d);
            acc[d] = 0.f;
        }
        float mx = -INFINITY, l = 0.f;
        for (int tt = t_lo; tt <= t_hi; tt++) {
            const int ct = tt - t_lo;
            for (int dy = -rs; dy <= rs; dy++) {
                const int cy = (q >> 3) + rs + dy;
                for (int dx = -rs; dx <= rs; dx++) {
                    const int c = (ct * hh + cy) * hw + (q & 7) + rs + dx;
                    lds_row<D4>(Ks + (size_t)c * HD + sub * D4, kv);
                    float s = 0.f;
#pragma unroll
                    for (int d = 0; d < D4; d++) s += qv[d] * kv[d];
                    s += __shfl_xor(s, 1, 64);
                    s += __shfl_xor(s, 2, 64);
                    s *= scale;
                    const float mn = fmaxf(mx, s);
                    const float corr = expf(mx - mn), p = expf(s - mn);
                    l = l * corr + p;
                    lds_row<D4>(Vs + (size_t)c * HD + sub * D4, kv);
#pragma unroll
                    for (int d = 0; d < D4; d++) acc[d] = acc[d] * corr + p * kv[d];
                    mx = mn;
                }
            }
        }
        if (live) {
            const float inv = 1.0f / l;
            T* o = out + qrow * dim + h * HD + sub * D4;
#pragma unroll
            for (int d = 0; d < D4; d++) {
                if constexpr (sizeof(T) == 4) o[d] = acc[d] * inv;
                else o[d] = f2bf(acc[d] * inv);
            }
        }
    }
}

// Dynamic LDS above 64 KB has to be allowed per kernel; done once per device
// at model creation (outside any stream capture).
void attn_cell_prepare() {
    const void* fns[] = {(const void*)k_attn_cell<8, float>,     (const void*)k_attn_cell<16, float>,
                         (const void*)k_attn_cell<32, float>,    (const void*)k_attn_cell<64, float>,
                         (const void*)k_attn_cell<8, uint16_t>,  (const void*)k_attn_cell<16, uint16_t>,
                         (const void*)k_attn_cell<32, uint16_t>, (const void*)k_attn_cell<64, uint16_t>};
    for (const void* f : fns) HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
}

size_t attn_cell_smem(int hd, int rs, int rt, int elem) {
    return 2ull * (2 * rt + 1) * (CT_Y + 2 * rs) * (CT_X + 2 * rs) * hd * elem;
}

void build_cell_tiles(const std::vector<Seg>& segs, std::vector<CellTile>& tiles) {
    tiles.clear();
    for (size_t i = 0; i < segs.size(); i++)
        for (int t = 0; t < segs[i].T; t++)
            for (int y = 0; y < segs[i].ny; y += CT_Y)
                for (int x = 0; x < segs[i].nx; x += CT_X) tiles.push_back(CellTile{(int)i, t, y, x});
}

template <class T>
static void launch(hipStream_t s, const T* qkv, T* out, int dim, int heads, int rs, int rt, const Seg* segs,
                   const CellTile* tiles, int n_tiles) {
    if (n_tiles <= 0) return;
    const int hd = dim / heads;
    const size_t smem = attn_cell_smem(hd, rs, rt, sizeof(T));
    if (smem > 160 * 1024) fail(J2K_ERR_LIMIT, "attention: stencil neighbourhood does not fit in LDS");
    dim3 grid(n_tiles), block(256);
    switch (hd) {
        case 8: hipLaunchKernelGGL((k_attn_cell<8, T>), grid, block, smem, s, qkv, out, dim, heads, rs, rt, segs, tiles); break;
        case 16: hipLaunchKernelGGL((k_attn_cell<16, T>), grid, block, smem, s, qkv, out, dim, heads, rs, rt, segs, tiles); break;
        case 32: hipLaunchKernelGGL((k_attn_cell<32, T>), grid, block, smem, s, qkv, out, dim, heads, rs, rt, segs, tiles); break;
        case 64: hipLaunchKernelGGL((k_attn_cell<64, T>), grid, block, smem, s, qkv, out, dim, heads, rs, rt, segs, tiles); break;
        default: fail(J2K_ERR_LIMIT, "attention: head size must be 8, 16, 32 or 64");
    }
    LAUNCHCHK("attention (cell tiles)");
}

void attention_cell_f32(hipStream_t s, const float* qkv, float* out, int dim, int heads, int rs, int rt,
                        const Seg* segs, const CellTile* tiles, int n_tiles) {
    launch<float>(s, qkv, out, dim, heads, rs, rt, segs, tiles, n_tiles);
}

void attention_cell_bf16(hipStream_t s, const uint16_t* qkv, uint16_t* out, int dim, int heads, int rs, int rt,
                         const Seg* segs, const CellTile* tiles, int n_tiles) {
    launch<uint16_t>(s, qkv, out, dim, heads, rs, rt, segs, tiles, n_tiles);
}

}  // namespace j2k
