// Stencil attention of the cell_st form with K/V staged in LDS.
//
// A block owns an 8 x 8 tile of cells in one time slice of one segment and
// loops over the heads. For each head it stages the K and V rows of the
// tile's neighbourhood (the tile plus a periodic halo of RS cells in y and x,
// over the clipped time window t-RT .. t+RT) into LDS once, so every key row
// is read from L2 once per tile instead of once per query. Four lanes share a
// query, each holding a quarter of the head's dimensions; the dot products
// are reduced with two lane shuffles. The softmax is two-pass over the
// scores kept in registers (max, then exp / sum / weighted V in one sweep),
// f32 throughout; keys are visited in the order time offset, y offset,
// x offset (the order of the generic kernel and of the oracle).
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

// N consecutive elements from LDS as f32, in 16-byte vectors.
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

__device__ __forceinline__ int wrap(int v, int n) {
    while (v < 0) v += n;
    while (v >= n) v -= n;
    return v;
}

template <int HD, int RS, int RT, class T>
__global__ __launch_bounds__(256) void k_attn_cell(const T* __restrict__ qkv, T* __restrict__ out, int dim, int heads,
                                                   const Seg* __restrict__ segs, const CellTile* __restrict__ tiles) {
    constexpr int D4 = HD / 4;               // dims per lane
    constexpr int VEC = 16 / sizeof(T);      // elements per 16-byte vector
    constexpr int VPR = HD / VEC;            // vectors per K (or V) row
    constexpr int HW = CT_X + 2 * RS, HH = CT_Y + 2 * RS, PLANE = HH * HW;
    constexpr int KMAX = (2 * RS + 1) * (2 * RS + 1) * (2 * RT + 1);
    extern __shared__ __align__(16) unsigned char smem[];
    const CellTile tl = tiles[blockIdx.x];
    const Seg sg = segs[tl.seg];
    const int t_lo = max(tl.t - RT, 0), t_hi = min(tl.t + RT, sg.T - 1), nt = t_hi - t_lo + 1;
    const int ncell = nt * PLANE;
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
        for (int v = tid; v < ncell * 2 * VPR; v += 256) {
            const int c = v / (2 * VPR), rem = v - c * (2 * VPR), which = rem / VPR, e = rem - which * VPR;
            const int ct = c / PLANE, cr = c - ct * PLANE, cy = cr / HW, cx = cr - cy * HW;
            const int y = wrap(tl.y0 - RS + cy, sg.ny), x = wrap(tl.x0 - RS + cx, sg.nx);
            const long row = (long)sg.base + (long)(t_lo + ct) * n + (long)y * sg.nx + x;
            const uint4 val = *reinterpret_cast<const uint4*>(qkv + row * ld + (1 + which) * dim + h * HD + e * VEC);
            *reinterpret_cast<uint4*>((which ? Vs : Ks) + (size_t)c * HD + e * VEC) = val;
        }
        __syncthreads();
        float qv[D4], kv[D4], acc[D4], s[KMAX];
#pragma unroll
        for (int d = 0; d < D4; d++) qv[d] = ldf<T>(qkv + qrow * ld + h * HD + sub * D4 + d);
        float mx = -INFINITY;
        int nk = 0;
#pragma unroll
        for (int ct = 0; ct < 2 * RT + 1; ct++) {
            if (ct >= nt) break;
#pragma unroll
            for (int dy = -RS; dy <= RS; dy++)
#pragma unroll
                for (int dx = -RS; dx <= RS; dx++) {
                    const int c = (ct * HH + (q >> 3) + RS + dy) * HW + (q & 7) + RS + dx;
                    lds_row<D4>(Ks + (size_t)c * HD + sub * D4, kv);
                    float dot = 0.f;
#pragma unroll
                    for (int d = 0; d < D4; d++) dot += qv[d] * kv[d];
                    dot += __shfl_xor(dot, 1, 64);
                    dot += __shfl_xor(dot, 2, 64);
                    const float sc = dot * scale;
                    s[ct * (2 * RS + 1) * (2 * RS + 1) + (dy + RS) * (2 * RS + 1) + dx + RS] = sc;
                    mx = fmaxf(mx, sc);
                    nk++;
                }
        }
#pragma unroll
        for (int d = 0; d < D4; d++) acc[d] = 0.f;
        float l = 0.f;
#pragma unroll
        for (int ct = 0; ct < 2 * RT + 1; ct++) {
            if (ct >= nt) break;
#pragma unroll
            for (int dy = -RS; dy <= RS; dy++)
#pragma unroll
                for (int dx = -RS; dx <= RS; dx++) {
                    const int c = (ct * HH + (q >> 3) + RS + dy) * HW + (q & 7) + RS + dx;
                    const float p = __expf(s[ct * (2 * RS + 1) * (2 * RS + 1) + (dy + RS) * (2 * RS + 1) + dx + RS] - mx);
                    l += p;
                    lds_row<D4>(Vs + (size_t)c * HD + sub * D4, kv);
#pragma unroll
                    for (int d = 0; d < D4; d++) acc[d] += p * kv[d];
                }
        }
        (void)nk;
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

namespace {

template <int HD, int RS, int RT, class T>
void* fn_ptr() {
    return (void*)k_attn_cell<HD, RS, RT, T>;
}

// Kernel for (head size, space radius, time radius); nullptr when not built.
template <class T>
const void* pick(int hd, int rs, int rt) {
#define J2K_ATTN_CASE(HD_, RS_, RT_) \
    if (hd == HD_ && rs == RS_ && rt == RT_) return fn_ptr<HD_, RS_, RT_, T>();
#define J2K_ATTN_HD(HD_)      \
    J2K_ATTN_CASE(HD_, 1, 0)  \
    J2K_ATTN_CASE(HD_, 1, 1)  \
    J2K_ATTN_CASE(HD_, 2, 0)  \
    J2K_ATTN_CASE(HD_, 2, 1)
    J2K_ATTN_HD(8)
    J2K_ATTN_HD(16)
    J2K_ATTN_HD(32)
    J2K_ATTN_HD(64)
#undef J2K_ATTN_HD
#undef J2K_ATTN_CASE
    return nullptr;
}

}  // namespace

bool attn_cell_ok(int hd, int rs, int rt, int elem) {
    const void* f = elem == 4 ? pick<float>(hd, rs, rt) : pick<uint16_t>(hd, rs, rt);
    return f != nullptr && attn_cell_smem(hd, rs, rt, elem) <= 160 * 1024;
}

// Dynamic LDS above 64 KB has to be allowed per kernel; done once per device
// at model creation (outside any stream capture).
void attn_cell_prepare() {
    for (int hd : {8, 16, 32, 64})
        for (int rs : {1, 2})
            for (int rt : {0, 1}) {
                HIPCHK(hipFuncSetAttribute(pick<float>(hd, rs, rt), hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
                HIPCHK(hipFuncSetAttribute(pick<uint16_t>(hd, rs, rt), hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            }
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
    const void* f = pick<T>(hd, rs, rt);
    const size_t smem = attn_cell_smem(hd, rs, rt, sizeof(T));
    if (!f || smem > 160 * 1024) fail(J2K_ERR_LIMIT, "attention: no LDS-tiled kernel for this head size / stencil");
    void* args[] = {(void*)&qkv, (void*)&out, (void*)&dim, (void*)&heads, (void*)&segs, (void*)&tiles};
    HIPCHK(hipLaunchKernel(f, dim3(n_tiles), dim3(256), args, smem, s));
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
