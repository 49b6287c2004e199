// Stencil attention of the cell_st form with K/V staged in LDS.
//
// A block owns an 8 x 8 tile of cells in two consecutive time slices of one
// segment (four lanes per query, 512 threads) and loops over the heads. For
// each head it stages the K and V rows of the tile's neighbourhood (the tile
// plus a periodic halo of RS cells in y and x, over the clipped time window
// t-RT .. t+1+RT) into LDS once, so every key row is read from L2 once per
// tile instead of once per query, and the two slices share the planes
// between them. Rows are loaded
// into registers one phase ahead of their use, so the L2 round trips overlap
// the arithmetic. Tiles are dealt to blocks XCD by XCD (xcd_remap):
// consecutive tiles (a band of tile rows through every slice), and so the
// time slices whose K/V rows they share, sit in one XCD's L2.
//
// Three staging modes, chosen by the autotuner (the arithmetic, and so the
// result, is the same in all):
//   KV_JOINT  K and V of a head sit in LDS together; one phase per head.
//   KV_SPLIT  K, then V, in one buffer of half the size: twice the blocks
//             per CU, twice the barriers.
//   KV_PLANE  one time plane of K or V at a time, double-buffered: a few KB
//             of LDS, one barrier per plane.
//   KV_PAIR2, KV_PAIR3  KV_SPLIT with two x-adjacent queries per lane, which
//             share their key rows in registers (k_attn_cell_pair).
//
// The four lanes of a query each read a quarter of every key row, as
// 16-byte chunks, in an order set by the query's x position (read_order).
// gfx950's LDS serves a 16-byte read to 16 lanes a cycle from 64 four-byte
// banks (and inside each 8-lane half, banks b and b + 16 of a 32-bank half
// must read the same 256-byte line); those 16 lanes are four queries at
// consecutive x, reading four consecutive staged rows. With 128-byte rows
// (f32, 32 dimensions) rows c and c + 2 lie on the same 32 banks: reading
// their chunks in opposite orders puts them on alternate banks, so each
// read takes one pass instead of two. 256-byte rows rotate the order over
// all four queries; shorter rows are conflict-free as they are.
//
// Four lanes share a query, each holding a quarter of the head's dimensions;
// the dot products are reduced with two DPP quad permutations. The softmax is
// two-pass over the scores kept in registers (max, then exp / sum / weighted
// V in one sweep), f32 throughout; keys are visited in the order time
// offset, y offset, x offset (the order of the generic kernel and of the
// oracle).
#include "common.h"
#include "kernels.h"

namespace j2k {

constexpr int CT_Y = 8, CT_X = 8, CT_T = 2;  // a tile: 8 x 8 cells of CT_T consecutive slices
constexpr int CNT = 256 * CT_T;               // threads of a block: four lanes per query

// A native 16-byte vector: copies of it stay in registers (HIP's uint4 is a
// struct whose copies lower to memcpy and pin arrays of it in scratch).
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// N consecutive elements as f32, in 16-byte vectors when N allows (the row
// then starts 16-byte aligned: head sizes and offsets are multiples of it).
template <int N>
__device__ __forceinline__ void row_f32(const float* p, float* o) {
    if constexpr (N % 4 == 0) {
#pragma unroll
        for (int i = 0; i < N; i += 4) {
            const float4 v = *reinterpret_cast<const float4*>(p + i);
            o[i] = v.x;
            o[i + 1] = v.y;
            o[i + 2] = v.z;
            o[i + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) o[i] = p[i];
    }
}
template <int N>
__device__ __forceinline__ void row_f32(const uint16_t* p, float* o) {
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
template <int N>
__device__ __forceinline__ void row_store(float* p, const float* v) {
    if constexpr (N % 4 == 0) {
#pragma unroll
        for (int i = 0; i < N; i += 4) *reinterpret_cast<float4*>(p + i) = make_float4(v[i], v[i + 1], v[i + 2], v[i + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) p[i] = v[i];
    }
}
template <int N>
__device__ __forceinline__ void row_store(uint16_t* p, const float* v) {
    if constexpr (N % 8 == 0) {
#pragma unroll
        for (int i = 0; i < N; i += 8) {
            uint32_t w[4];
#pragma unroll
            for (int j = 0; j < 4; j++) w[j] = (uint32_t)f2bf(v[i + 2 * j]) | ((uint32_t)f2bf(v[i + 2 * j + 1]) << 16);
            *reinterpret_cast<uint4*>(p + i) = make_uint4(w[0], w[1], w[2], w[3]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) p[i] = f2bf(v[i]);
    }
}

__device__ __forceinline__ int wrap(int v, int n) {
    while (v < 0) v += n;
    while (v >= n) v -= n;
    return v;
}

// The chunk order of a query at tile column qx (its lane's quarter of a
// row is LPC chunks; the lane reads chunk i ^ read_order at step i).
template <int HD, class T>
__device__ __forceinline__ int read_order(int qx) {
    constexpr int R = HD * (int)sizeof(T);
    if constexpr (R == 128) return (qx >> 1) & 1;
    else if constexpr (R == 256) return qx & 3;
    else return 0;
}

// The lane's quarter (sub) of a row as D4 f32 values, its LPC 16-byte
// chunks in the order g: value i*VEC + k is element k of chunk i ^ g. The
// query, the keys and the values of one lane all use its order, so a dot
// product pairs the same dimensions; its terms are added in that order.
template <int HD, class T>
__device__ __forceinline__ void row_quarter(const T* row, int g, int sub, float* o) {
    constexpr int D4 = HD / 4, VEC = 16 / (int)sizeof(T), LPC = D4 / VEC;
    if constexpr (D4 % VEC == 0 && LPC >= 2) {
#pragma unroll
        for (int i = 0; i < LPC; i++) row_f32<VEC>(row + (sub * LPC + (i ^ g)) * VEC, o + i * VEC);
    } else {
        row_f32<D4>(row + sub * D4, o);
    }
}
// Store of D4 values held in the order g (row_quarter) to their places.
template <int HD, class T>
__device__ __forceinline__ void row_store_quarter(T* row, int g, int sub, const float* v) {
    constexpr int D4 = HD / 4, VEC = 16 / (int)sizeof(T), LPC = D4 / VEC;
    if constexpr (D4 % VEC == 0 && LPC >= 2) {
#pragma unroll
        for (int i = 0; i < LPC; i++) row_store<VEC>(row + (sub * LPC + (i ^ g)) * VEC, v + i * VEC);
    } else {
        row_store<D4>(row + sub * D4, v);
    }
}

// Waves per SIMD for a block of CNT threads: two blocks per CU when the LDS
// footprint allows it (up to 80 KB) and a lane's scores and q / k /
// accumulator quarters fit half the register file: a 27-key stencil, or a
// wider one whose scores are kept a quarter per lane (up to 48 registers of
// both); else one block, whose register file then holds twice as much.
template <int HD, int RS, int RT, class T, int MODE>
constexpr int cell_occupancy() {
    constexpr int KMAX = (2 * RS + 1) * (2 * RS + 1) * (2 * RT + 1);
    return (MODE == KV_JOINT ? 2 : 1) * (2 * RT + CT_T) * (CT_Y + 2 * RS) * (CT_X + 2 * RS) * HD * (int)sizeof(T) >
                           80 * 1024 ||
                   (KMAX > 32 && (KMAX + 3) / 4 + 3 * HD / 4 > 48)
               ? CNT / 256
               : 2 * CNT / 256;
}

template <int HD, int RS, int RT, class T, int MODE>
__global__ __launch_bounds__(CNT, (cell_occupancy<HD, RS, RT, T, MODE>())) void k_attn_cell(
    const T* __restrict__ qkv, T* __restrict__ out, int dim, int heads, const Seg* __restrict__ segs,
    const CellTile* __restrict__ tiles) {
    constexpr int D4 = HD / 4;               // dims per lane
    constexpr int VEC = 16 / sizeof(T);      // elements per 16-byte vector
    constexpr int VPR = HD / VEC;            // vectors per K (or V) row
    constexpr int HW = CT_X + 2 * RS, HH = CT_Y + 2 * RS, PLANE = HH * HW;
    constexpr int KS = (2 * RS + 1) * (2 * RS + 1), KMAX = KS * (2 * RT + 1);
    // Above 32 keys, where that lets two blocks share a CU, the scores are
    // kept per group of four keys, one a lane (quad_scatter4): a quarter of
    // the registers, the same values.
    constexpr int OCC = cell_occupancy<HD, RS, RT, T, MODE>();
    constexpr bool GROUPED = KMAX > 32 && OCC > CNT / 256;
    constexpr int NG = (KMAX + 3) / 4, NS = GROUPED ? NG : KMAX;  // score registers a lane
    constexpr int ROWS = MODE == KV_JOINT ? 2 : 1;  // K/V rows per cell in one staging phase
    constexpr int NV = ((2 * RT + CT_T) * PLANE * ROWS * VPR + CNT - 1) / CNT;  // staged vectors per thread
    extern __shared__ __align__(16) unsigned char smem[];
    const CellTile tl = tiles[xcd_remap(blockIdx.x, gridDim.x)];
    const Seg sg = segs[tl.seg];
    // planes of the tile's slices and their time neighbours
    const int t_lo = max(tl.t - RT, 0), t_hi = min(tl.t + CT_T - 1 + RT, sg.T - 1), nt = t_hi - t_lo + 1;
    const int ncell = nt * PLANE, nvec = ncell * ROWS * VPR;
    T* Ks = reinterpret_cast<T*>(smem);
    T* Vs = MODE == KV_JOINT ? Ks + (size_t)ncell * HD : Ks;
    const long ld = 3L * dim;
    const int n = sg.ny * sg.nx;
    const int tid = threadIdx.x;

    // Vector v of a staging phase holds elements [e*VEC, e*VEC+VEC) of the K
    // (which = 0) or V row of neighbourhood cell c: v = (c*ROWS + which)*VPR
    // + e (KV_JOINT), v = c*VPR + e with `which` set by the phase (KV_SPLIT).
    // Vectors past the neighbourhood (v >= nvec) load the segment's first row
    // and are never stored, so every load is unconditional.
    auto row_of = [&](int v) {
        if (v >= nvec) return sg.base;
        const int c = v / (ROWS * VPR), ct = c / PLANE, cr = c - ct * PLANE, cy = cr / HW, cx = cr - cy * HW;
        const int y = wrap(tl.y0 - RS + cy, sg.ny), x = wrap(tl.x0 - RS + cx, sg.nx);
        return sg.base + (t_lo + ct) * n + y * sg.nx + x;
    };
    // The first PF vectors of each thread are loaded into registers one phase
    // ahead; the rest are loaded when the phase is staged, SB at a time. PF
    // is sized so that the prefetch (5 registers a vector), the scores and
    // the q / k / accumulator rows fit the register file without spilling.
    // The grouped (wider) stencils prefetch nothing and load a phase's rows
    // in one batch: their second block per CU hides that latency, and the
    // registers a prefetch would hold through the arithmetic do not fit
    // beside it (a 75-key, 32-dimension head: 551 us per 262,144-token
    // launch, against 597-714 with two to four vectors ahead and 764 with
    // the scores of every key in registers at one block per CU).
    constexpr int PF_FIT = (512 / OCC * 7 / 10 - NS - 3 * D4) / 5;
    constexpr int PF = GROUPED ? 0 : NV < PF_FIT ? NV : (PF_FIT < 4 ? 4 : PF_FIT);
    constexpr int SB = GROUPED ? NV : 8;
    int grow[PF > 0 ? PF : 1];
#pragma unroll
    for (int i = 0; i < PF; i++) grow[i] = row_of(tid + CNT * i);
    // Loads of phase (head h, rows `w`: 0 = K, 1 = V, -1 = both) and their
    // stores into LDS; written out in place rather than as lambdas so that
    // `pre` stays in registers.
#define J2K_WHICH(v, w) ((w) < 0 ? ((v) / VPR) & 1 : (w))
#define J2K_FETCH(h, w)                                                                                            \
    _Pragma("unroll") for (int i = 0; i < PF; i++) {                                                               \
        const int v = tid + CNT * i, e = v % VPR;                                                                  \
        pre[i] = *reinterpret_cast<const u32x4*>(qkv + (long)grow[i] * ld + (1 + J2K_WHICH(v, w)) * dim + (h) * HD + \
                                                 e * VEC);                                                         \
    }
#define J2K_STAGE(h, w)                                                                                            \
    {                                                                                                              \
        _Pragma("unroll") for (int i = 0; i < PF; i++) {                                                           \
            const int v = tid + CNT * i, c = v / (ROWS * VPR), e = v % VPR;                                        \
            if (v < nvec) *reinterpret_cast<u32x4*>((J2K_WHICH(v, w) ? Vs : Ks) + (size_t)c * HD + e * VEC) = pre[i]; \
        }                                                                                                          \
        _Pragma("unroll") for (int i0 = PF; i0 < NV; i0 += SB) {                                                   \
            u32x4 tmp[SB];                                                                                         \
            _Pragma("unroll") for (int j = 0; j < SB; j++) {                                                       \
                const int i = i0 + j, v = tid + CNT * i, e = v % VPR;                                              \
                if (i < NV)                                                                                        \
                    tmp[j] = *reinterpret_cast<const u32x4*>(qkv + (long)row_of(v) * ld + (1 + J2K_WHICH(v, w)) * dim + \
                                                             (h) * HD + e * VEC);                                  \
            }                                                                                                      \
            _Pragma("unroll") for (int j = 0; j < SB; j++) {                                                       \
                const int i = i0 + j, v = tid + CNT * i, c = v / (ROWS * VPR), e = v % VPR;                        \
                if (i < NV && v < nvec)                                                                            \
                    *reinterpret_cast<u32x4*>((J2K_WHICH(v, w) ? Vs : Ks) + (size_t)c * HD + e * VEC) = tmp[j];    \
            }                                                                                                      \
        }                                                                                                          \
    }
    u32x4 pre[PF > 0 ? PF : 1];
    // the lane's query: cell q of the tile in slice tq (a slice past the
    // segment's end repeats the tile's first and stores nothing); its keys
    // are planes poff .. poff + ntq - 1 of the staged ones
    const int q = (tid >> 2) & 63, sub = tid & 3, ts = tid >> 8, g = read_order<HD, T>(q & 7);
    const int qy = tl.y0 + (q >> 3), qx = tl.x0 + (q & 7);
    const bool live = qy < sg.ny && qx < sg.nx && tl.t + ts < sg.T;
    const int tq = tl.t + ts < sg.T ? tl.t + ts : tl.t;
    const int tq_lo = max(tq - RT, 0), ntq = min(tq + RT, sg.T - 1) - tq_lo + 1, poff = tq_lo - t_lo;
    const long qrow = (long)sg.base + (long)tq * n + (long)(live ? qy : 0) * sg.nx + (live ? qx : 0);
    const float scale = 1.0f / sqrtf((float)HD);
    // a wave's 64 lanes are 16 queries of one slice, so their time window
    // is one: the grouped key loops branch on it as a scalar
    const int ntq_w = __builtin_amdgcn_readfirstlane(ntq);
    // the query's first key row (time offset 0 of its window, y and x
    // offsets -RS): key (ct, dy, dx) lies a compile-time distance past it
    const int k0 = (poff * HH + (q >> 3)) * HW + (q & 7);
    const T* const kb = Ks + k0 * HD;
    const T* const vb = Vs + k0 * HD;

    if (MODE == KV_JOINT) {
        J2K_FETCH(0, -1)
    } else {
        J2K_FETCH(0, 0)
    }
    for (int h = 0; h < heads; h++) {
        __syncthreads();  // the previous phase's rows are no longer read
        if (MODE == KV_JOINT) {
            J2K_STAGE(h, -1)
        } else {
            J2K_STAGE(h, 0)
        }
        __syncthreads();
        if (MODE == KV_JOINT) {
            if (h + 1 < heads) {
                J2K_FETCH(h + 1, -1)  // in flight while this head is computed
            }
        } else {
            J2K_FETCH(h, 1)  // V of this head, in flight during the scores
        }
        float qv[D4], kv[D4], acc[D4], s[NS];
        row_quarter<HD, T>(qkv + qrow * ld + h * HD, g, sub, qv);
        float mx = -INFINITY;
        if constexpr (GROUPED) {
#pragma unroll
            for (int gi = 0; gi < NG; gi++) {
                float pd[4];  // the lane's partial dot products of the group's four keys
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    constexpr int W = 2 * RS + 1;
                    const int k = 4 * gi + j, ct = k / KS, dy = (k % KS) / W, dx = k % W;
                    pd[j] = 0.f;
                    if (k < KMAX && ct < ntq_w) {
                        row_quarter<HD, T>(kb + ((ct * HH + dy) * HW + dx) * HD, g, sub, kv);
                        float dot = 0.f;
#pragma unroll
                        for (int d = 0; d < D4; d++) dot += qv[d] * kv[d];
                        pd[j] = dot;
                    }
                }
                const float dot = quad_scatter4(pd, sub);  // key 4 gi + sub
                const int k = 4 * gi + sub;
                const float sc = k < KMAX && k / KS < ntq_w ? dot * scale : -INFINITY;
                s[gi] = sc;
                mx = fmaxf(mx, sc);
            }
            mx = fmaxf(mx, quad_xor1(mx));
            mx = fmaxf(mx, quad_xor2(mx));
        } else
#pragma unroll
        for (int ct = 0; ct < 2 * RT + 1; ct++) {
            if (ct >= ntq) break;
#pragma unroll
            for (int dy = -RS; dy <= RS; dy++)
#pragma unroll
                for (int dx = -RS; dx <= RS; dx++) {
                    row_quarter<HD, T>(kb + ((ct * HH + RS + dy) * HW + RS + dx) * HD, g, sub, kv);
                    float dot = 0.f;
#pragma unroll
                    for (int d = 0; d < D4; d++) dot += qv[d] * kv[d];
                    dot += quad_xor1(dot);
                    dot += quad_xor2(dot);
                    const float sc = dot * scale;
                    s[ct * KS + (dy + RS) * (2 * RS + 1) + dx + RS] = sc;
                    mx = fmaxf(mx, sc);
                }
        }
        if (MODE == KV_SPLIT) {
            __syncthreads();  // every wave's scores are done with K
            J2K_STAGE(h, 1)
            __syncthreads();
            if (h + 1 < heads) {
                J2K_FETCH(h + 1, 0)  // next head's K, in flight during the weighting
            }
        }
#pragma unroll
        for (int d = 0; d < D4; d++) acc[d] = 0.f;
        float l = 0.f;
        if constexpr (GROUPED) {
#pragma unroll
            for (int gi = 0; gi < NG; gi++) {
                const float po = __expf(s[gi] - mx);  // the weight of the lane's own key of the group
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    constexpr int W = 2 * RS + 1;
                    const int k = 4 * gi + j, ct = k / KS, dy = (k % KS) / W, dx = k % W;
                    if (k < KMAX && ct < ntq_w) {
                        const float p = quad_bcast(po, j);
                        l += p;
                        row_quarter<HD, T>(vb + ((ct * HH + dy) * HW + dx) * HD, g, sub, kv);
#pragma unroll
                        for (int d = 0; d < D4; d++) acc[d] += p * kv[d];
                    }
                }
            }
        } else
#pragma unroll
        for (int ct = 0; ct < 2 * RT + 1; ct++) {
            if (ct >= ntq) break;
#pragma unroll
            for (int dy = -RS; dy <= RS; dy++)
#pragma unroll
                for (int dx = -RS; dx <= RS; dx++) {
                    const float p = __expf(s[ct * KS + (dy + RS) * (2 * RS + 1) + dx + RS] - mx);
                    l += p;
                    row_quarter<HD, T>(vb + ((ct * HH + RS + dy) * HW + RS + dx) * HD, g, sub, kv);
#pragma unroll
                    for (int d = 0; d < D4; d++) acc[d] += p * kv[d];
                }
        }
        if (live) {
            const float inv = 1.0f / l;
#pragma unroll
            for (int d = 0; d < D4; d++) acc[d] *= inv;
            row_store_quarter<HD, T>(out + qrow * dim + h * HD, g, sub, acc);
        }
    }
#undef J2K_STAGE
#undef J2K_FETCH
#undef J2K_WHICH
}

// KV_PLANE: one time plane of K, then of V, at a time. Per head the phases
// are K of slices t_lo .. t_hi (scores of the queries whose window holds the
// slice), then V of the same slices (weights),
// each plane (tile + halo, one head) double-buffered in LDS with the next
// phase's rows already in flight in registers: one barrier per phase and a
// small LDS footprint, so more blocks share a CU. The arithmetic, and its
// order, is that of the other modes.
// Four waves per SIMD (two blocks) for the 27-key stencils; the scores of
// wider ones need the registers of two.
template <int HD, int RS, int RT, class T>
__global__ __launch_bounds__(CNT, ((2 * RS + 1) * (2 * RS + 1) * (2 * RT + 1) > 32 || HD > 32 ? 2 : 4)) void k_attn_cell_plane(const T* __restrict__ qkv, T* __restrict__ out, int dim,
                                                           int heads, const Seg* __restrict__ segs,
                                                           const CellTile* __restrict__ tiles) {
    constexpr int D4 = HD / 4, VEC = 16 / sizeof(T), VPR = HD / VEC;
    constexpr int HW = CT_X + 2 * RS, HH = CT_Y + 2 * RS, PLANE = HH * HW;
    constexpr int KS = (2 * RS + 1) * (2 * RS + 1), KMAX = KS * (2 * RT + 1);
    constexpr int NV = (PLANE * VPR + CNT - 1) / CNT;  // staged vectors per thread per plane
    extern __shared__ __align__(16) unsigned char smem[];  // two planes
    T* const planes = reinterpret_cast<T*>(smem);
    const CellTile tl = tiles[xcd_remap(blockIdx.x, gridDim.x)];
    const Seg sg = segs[tl.seg];
    const int t_lo = max(tl.t - RT, 0), t_hi = min(tl.t + CT_T - 1 + RT, sg.T - 1), nt = t_hi - t_lo + 1;
    const long ld = 3L * dim;
    const int n = sg.ny * sg.nx;
    const int tid = threadIdx.x;

    // vector v of a plane: elements [e*VEC, e*VEC+VEC) of neighbourhood cell
    // c = v / VPR; its offset in a slice (the same for every slice and head).
    // Vectors past the plane load the tile's own slice start and are dropped.
    int coff[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) {
        const int v = tid + CNT * i, c = v / VPR, cy = c / HW, cx = c - cy * HW;
        coff[i] = v < PLANE * VPR ? wrap(tl.y0 - RS + cy, sg.ny) * sg.nx + wrap(tl.x0 - RS + cx, sg.nx) : 0;
    }
    u32x4 pre[NV];
    // phase f of a head: K of slice t_lo + f for f < nt, then V of slice t_lo + f - nt
#define J2K_PLANE_FETCH(h, f)                                                                                     \
    {                                                                                                            \
        const int which_ = (f) < nt ? 0 : 1, t_ = t_lo + ((f) < nt ? (f) : (f) - nt);                            \
        const T* src_ = qkv + ((long)sg.base + (long)t_ * n) * ld + (1 + which_) * dim + (h) * HD;              \
        _Pragma("unroll") for (int i = 0; i < NV; i++) {                                                          \
            const int v = tid + CNT * i, e = v % VPR;                                                            \
            pre[i] = *reinterpret_cast<const u32x4*>(src_ + (long)coff[i] * ld + e * VEC);                       \
        }                                                                                                        \
    }

    // the lane's query, as in k_attn_cell: its key planes are poff .. poff + ntq - 1
    const int q = (tid >> 2) & 63, sub = tid & 3, ts = tid >> 8, g = read_order<HD, T>(q & 7);
    const int qy = tl.y0 + (q >> 3), qx = tl.x0 + (q & 7);
    const bool live = qy < sg.ny && qx < sg.nx && tl.t + ts < sg.T;
    const int tq = tl.t + ts < sg.T ? tl.t + ts : tl.t;
    const int tq_lo = max(tq - RT, 0), ntq = min(tq + RT, sg.T - 1) - tq_lo + 1, poff = tq_lo - t_lo;
    const long qrow = (long)sg.base + (long)tq * n + (long)(live ? qy : 0) * sg.nx + (live ? qx : 0);
    const float scale = 1.0f / sqrtf((float)HD);
    const int phases = 2 * nt;
    const int k0 = (q >> 3) * HW + (q & 7);  // the query's key at y and x offsets -RS in a plane

    J2K_PLANE_FETCH(0, 0)
    int b = 0;
    for (int h = 0; h < heads; h++) {
        float qv[D4], kv[D4], acc[D4], s[KMAX];
        row_quarter<HD, T>(qkv + qrow * ld + h * HD, g, sub, qv);
        float mx = -INFINITY, l = 0.f;
#pragma unroll
        for (int d = 0; d < D4; d++) acc[d] = 0.f;
        for (int f = 0; f < phases; f++, b ^= 1) {
            T* P = planes + (size_t)b * PLANE * HD;
#pragma unroll
            for (int i = 0; i < NV; i++) {
                const int v = tid + CNT * i, c = v / VPR, e = v % VPR;
                if (v < PLANE * VPR) *reinterpret_cast<u32x4*>(P + (size_t)c * HD + e * VEC) = pre[i];
            }
            __syncthreads();  // the plane is in; every thread is past the phase before the last
            if (f + 1 < phases) {
                J2K_PLANE_FETCH(h, f + 1)
            } else if (h + 1 < heads) {
                J2K_PLANE_FETCH(h + 1, 0)
            }
            if (f < nt) {
                // scores of slice t_lo + f
#pragma unroll
                for (int ct = 0; ct < 2 * RT + 1; ct++) {
                    if (ct != f - poff || ct >= ntq) continue;
#pragma unroll
                    for (int dy = -RS; dy <= RS; dy++)
#pragma unroll
                        for (int dx = -RS; dx <= RS; dx++) {
                            row_quarter<HD, T>(P + (k0 + (RS + dy) * HW + RS + dx) * HD, g, sub, kv);
                            float dot = 0.f;
#pragma unroll
                            for (int d = 0; d < D4; d++) dot += qv[d] * kv[d];
                            dot += quad_xor1(dot);
                            dot += quad_xor2(dot);
                            const float sc = dot * scale;
                            s[ct * KS + (dy + RS) * (2 * RS + 1) + dx + RS] = sc;
                            mx = fmaxf(mx, sc);
                        }
                }
            } else {
                // weights of slice t_lo + f - nt
#pragma unroll
                for (int ct = 0; ct < 2 * RT + 1; ct++) {
                    if (ct != f - nt - poff || ct >= ntq) continue;
#pragma unroll
                    for (int dy = -RS; dy <= RS; dy++)
#pragma unroll
                        for (int dx = -RS; dx <= RS; dx++) {
                            const float p = __expf(s[ct * KS + (dy + RS) * (2 * RS + 1) + dx + RS] - mx);
                            l += p;
                            row_quarter<HD, T>(P + (k0 + (RS + dy) * HW + RS + dx) * HD, g, sub, kv);
#pragma unroll
                            for (int d = 0; d < D4; d++) acc[d] += p * kv[d];
                        }
                }
            }
        }
        if (live) {
            const float inv = 1.0f / l;
#pragma unroll
            for (int d = 0; d < D4; d++) acc[d] *= inv;
            row_store_quarter<HD, T>(out + qrow * dim + h * HD, g, sub, acc);
        }
    }
#undef J2K_PLANE_FETCH
}

// KV_PAIR2 / KV_PAIR3: the KV_SPLIT staging with two y-adjacent queries per
// lane (four lanes per pair, 256 threads per 8 x 8 x 2 tile). The pair's
// stencils overlap: for each time offset the lane reads 2 RS + 2 rows of
// 2 RS + 1 keys from LDS instead of 2 (2 RS + 1) rows, so the LDS reads fall
// by a third (27-key stencil). The 16 lanes of a read cycle are four pairs
// at consecutive x, which read four consecutive staged rows, as in the other
// modes. Keys are visited by time offset, key row, x offset, which for each
// query is its own time, y, x order: each query's dot products, softmax and
// weighting are those of the other modes, term for term and in the same
// order. OCC blocks share a CU (2: the whole staging phase is
// prefetched in registers; 3: part of it is loaded when it is staged).
template <int HD, int RS, int RT, class T, int OCC>
__global__ __launch_bounds__(256, OCC) void k_attn_cell_pair(const T* __restrict__ qkv, T* __restrict__ out, int dim,
                                                             int heads, const Seg* __restrict__ segs,
                                                             const CellTile* __restrict__ tiles) {
    constexpr int NT = 256;
    constexpr int D4 = HD / 4, VEC = 16 / sizeof(T), VPR = HD / VEC;
    constexpr int HW = CT_X + 2 * RS, HH = CT_Y + 2 * RS, PLANE = HH * HW;
    constexpr int KW = 2 * RS + 1, KS = KW * KW, KMAX = KS * (2 * RT + 1);
    constexpr int NV = ((2 * RT + CT_T) * PLANE * VPR + NT - 1) / NT;  // staged vectors per thread
    extern __shared__ __align__(16) unsigned char smem[];
    const CellTile tl = tiles[xcd_remap(blockIdx.x, gridDim.x)];
    const Seg sg = segs[tl.seg];
    const int t_lo = max(tl.t - RT, 0), t_hi = min(tl.t + CT_T - 1 + RT, sg.T - 1), nt = t_hi - t_lo + 1;
    const int ncell = nt * PLANE, nvec = ncell * VPR;
    T* const Ks = reinterpret_cast<T*>(smem);  // K, then V, of one head
    const long ld = 3L * dim;
    const int n = sg.ny * sg.nx;
    const int tid = threadIdx.x;
    auto row_of = [&](int v) {
        if (v >= nvec) return sg.base;
        const int c = v / VPR, ct = c / PLANE, cr = c - ct * PLANE, cy = cr / HW, cx = cr - cy * HW;
        const int y = wrap(tl.y0 - RS + cy, sg.ny), x = wrap(tl.x0 - RS + cx, sg.nx);
        return sg.base + (t_lo + ct) * n + y * sg.nx + x;
    };
    constexpr int PF_FIT = (512 / OCC * (OCC > 2 ? 6 : 7) / 10 - 2 * KMAX - 5 * D4) / 5;
    constexpr int PF = NV < PF_FIT ? NV : (PF_FIT < 4 ? 4 : PF_FIT);
    int grow[PF];
#pragma unroll
    for (int i = 0; i < PF; i++) grow[i] = row_of(tid + NT * i);
    u32x4 pre[PF];
#define J2K_PAIR_FETCH(h, w)                                                                                      \
    _Pragma("unroll") for (int i = 0; i < PF; i++) {                                                              \
        const int v = tid + NT * i, e = v % VPR;                                                                  \
        pre[i] = *reinterpret_cast<const u32x4*>(qkv + (long)grow[i] * ld + (1 + (w)) * dim + (h) * HD + e * VEC);   \
    }
#define J2K_PAIR_STAGE(h, w)                                                                                      \
    {                                                                                                             \
        _Pragma("unroll") for (int i = 0; i < PF; i++) {                                                          \
            const int v = tid + NT * i;                                                                           \
            if (v < nvec) *reinterpret_cast<u32x4*>(Ks + (size_t)v * VEC) = pre[i];                              \
        }                                                                                                         \
        _Pragma("unroll") for (int i0 = PF; i0 < NV; i0 += 8) {                                                   \
            u32x4 tmp[8];                                                                                         \
            _Pragma("unroll") for (int j = 0; j < 8; j++) {                                                       \
                const int i = i0 + j, v = tid + NT * i, e = v % VPR;                                              \
                if (i < NV)                                                                                       \
                    tmp[j] = *reinterpret_cast<const u32x4*>(qkv + (long)row_of(v) * ld + (1 + (w)) * dim + (h) * HD + \
                                                             e * VEC);                                            \
            }                                                                                                     \
            _Pragma("unroll") for (int j = 0; j < 8; j++) {                                                       \
                const int i = i0 + j, v = tid + NT * i;                                                           \
                if (i < NV && v < nvec) *reinterpret_cast<u32x4*>(Ks + (size_t)v * VEC) = tmp[j];                 \
            }                                                                                                     \
        }                                                                                                         \
    }
    // the lane's queries: cells (qyl, qxl) and (qyl + 1, qxl) of the tile in
    // slice tq (a slice past the segment's end repeats the tile's first and
    // stores nothing); keys are planes poff .. poff + ntq - 1 of the staged ones
    const int pr = (tid >> 2) & 31, sub = tid & 3, ts = tid >> 7;
    const int qyl = (pr >> 3) * 2, qxl = pr & 7, g = read_order<HD, T>(qxl);
    const int qy0 = tl.y0 + qyl, qx = tl.x0 + qxl;
    const bool tlive = qx < sg.nx && tl.t + ts < sg.T;
    const bool live0 = tlive && qy0 < sg.ny, live1 = tlive && qy0 + 1 < sg.ny;
    const int tq = tl.t + ts < sg.T ? tl.t + ts : tl.t;
    const int tq_lo = max(tq - RT, 0), ntq = min(tq + RT, sg.T - 1) - tq_lo + 1, poff = tq_lo - t_lo;
    const long slice0 = (long)sg.base + (long)tq * n;
    const long qrow0 = slice0 + (live0 ? (long)qy0 * sg.nx + qx : 0);
    const long qrow1 = slice0 + (live1 ? (long)(qy0 + 1) * sg.nx + qx : 0);
    const float scale = 1.0f / sqrtf((float)HD);
    // query 0's first key row (time offset 0 of its window, y and x offsets
    // -RS): key row e, x offset dx of time offset ct lies a compile-time
    // distance past it
    const T* const kb = Ks + ((poff * HH + qyl) * HW + qxl) * HD;

    J2K_PAIR_FETCH(0, 0)
    for (int h = 0; h < heads; h++) {
        __syncthreads();  // the previous phase's rows are no longer read
        J2K_PAIR_STAGE(h, 0)
        __syncthreads();
        J2K_PAIR_FETCH(h, 1)  // V of this head, in flight during the scores
        float q0[D4], q1[D4], kv[D4], a0[D4], a1[D4], s0[KMAX], s1[KMAX];
        row_quarter<HD, T>(qkv + qrow0 * ld + h * HD, g, sub, q0);
        row_quarter<HD, T>(qkv + qrow1 * ld + h * HD, g, sub, q1);
        float mx0 = -INFINITY, mx1 = -INFINITY;
#pragma unroll
        for (int ct = 0; ct < 2 * RT + 1; ct++) {
            if (ct >= ntq) break;
#pragma unroll
            for (int e = 0; e <= 2 * RS + 1; e++)
#pragma unroll
                for (int dx = -RS; dx <= RS; dx++) {
                    // key row y offset e - RS from query 0 and e - RS - 1 from query 1
                    row_quarter<HD, T>(kb + ((ct * HH + e) * HW + RS + dx) * HD, g, sub, kv);
                    if (e <= 2 * RS) {
                        float dot = 0.f;
#pragma unroll
                        for (int d = 0; d < D4; d++) dot += q0[d] * kv[d];
                        dot += quad_xor1(dot);
                        dot += quad_xor2(dot);
                        const float sc = dot * scale;
                        s0[ct * KS + e * KW + dx + RS] = sc;
                        mx0 = fmaxf(mx0, sc);
                    }
                    if (e >= 1) {
                        float dot = 0.f;
#pragma unroll
                        for (int d = 0; d < D4; d++) dot += q1[d] * kv[d];
                        dot += quad_xor1(dot);
                        dot += quad_xor2(dot);
                        const float sc = dot * scale;
                        s1[ct * KS + (e - 1) * KW + dx + RS] = sc;
                        mx1 = fmaxf(mx1, sc);
                    }
                }
        }
        __syncthreads();  // every wave's scores are done with K
        J2K_PAIR_STAGE(h, 1)
        __syncthreads();
        if (h + 1 < heads) {
            J2K_PAIR_FETCH(h + 1, 0)  // next head's K, in flight during the weighting
        }
#pragma unroll
        for (int d = 0; d < D4; d++) {
            a0[d] = 0.f;
            a1[d] = 0.f;
        }
        float l0 = 0.f, l1 = 0.f;
#pragma unroll
        for (int ct = 0; ct < 2 * RT + 1; ct++) {
            if (ct >= ntq) break;
#pragma unroll
            for (int e = 0; e <= 2 * RS + 1; e++)
#pragma unroll
                for (int dx = -RS; dx <= RS; dx++) {
                    row_quarter<HD, T>(kb + ((ct * HH + e) * HW + RS + dx) * HD, g, sub, kv);
                    if (e <= 2 * RS) {
                        const float p = __expf(s0[ct * KS + e * KW + dx + RS] - mx0);
                        l0 += p;
#pragma unroll
                        for (int d = 0; d < D4; d++) a0[d] += p * kv[d];
                    }
                    if (e >= 1) {
                        const float p = __expf(s1[ct * KS + (e - 1) * KW + dx + RS] - mx1);
                        l1 += p;
#pragma unroll
                        for (int d = 0; d < D4; d++) a1[d] += p * kv[d];
                    }
                }
        }
        if (live0) {
            const float inv = 1.0f / l0;
#pragma unroll
            for (int d = 0; d < D4; d++) a0[d] *= inv;
            row_store_quarter<HD, T>(out + qrow0 * dim + h * HD, g, sub, a0);
        }
        if (live1) {
            const float inv = 1.0f / l1;
#pragma unroll
            for (int d = 0; d < D4; d++) a1[d] *= inv;
            row_store_quarter<HD, T>(out + qrow1 * dim + h * HD, g, sub, a1);
        }
    }
#undef J2K_PAIR_STAGE
#undef J2K_PAIR_FETCH
}

namespace {

template <int HD, int RS, int RT, class T>
const void* fn_ptr(int mode) {
    switch (mode) {
        case KV_JOINT: return (const void*)k_attn_cell<HD, RS, RT, T, KV_JOINT>;
        case KV_SPLIT: return (const void*)k_attn_cell<HD, RS, RT, T, KV_SPLIT>;
        case KV_PLANE: return (const void*)k_attn_cell_plane<HD, RS, RT, T>;
        case KV_PAIR2: return (const void*)k_attn_cell_pair<HD, RS, RT, T, 2>;
        case KV_PAIR3: return (const void*)k_attn_cell_pair<HD, RS, RT, T, 3>;
        default: return nullptr;
    }
}

// Kernel for (head size, space radius, time radius, mode); nullptr when not built.
template <class T>
const void* pick(int hd, int rs, int rt, int mode) {
#define J2K_ATTN_CASE(HD_, RS_, RT_) \
    if (hd == HD_ && rs == RS_ && rt == RT_) return fn_ptr<HD_, RS_, RT_, T>(mode);
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

bool attn_cell_ok(int hd, int rs, int rt, int elem, int mode) {
    const void* f = elem == 4 ? pick<float>(hd, rs, rt, mode) : pick<uint16_t>(hd, rs, rt, mode);
    return f != nullptr && attn_cell_smem(hd, rs, rt, elem, mode) <= 160 * 1024;
}

// Dynamic LDS above 64 KB has to be allowed per kernel; done once per device
// at model creation (outside any stream capture).
void attn_cell_prepare() {
    for (int hd : {8, 16, 32, 64})
        for (int rs : {1, 2})
            for (int rt : {0, 1})
                for (int mode = 0; mode < KV_MODES; mode++) {
                    HIPCHK(hipFuncSetAttribute(pick<float>(hd, rs, rt, mode),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
                    HIPCHK(hipFuncSetAttribute(pick<uint16_t>(hd, rs, rt, mode),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
                }
}

size_t attn_cell_smem(int hd, int rs, int rt, int elem, int mode) {
    if (mode == KV_PLANE) return 2ull * (CT_Y + 2 * rs) * (CT_X + 2 * rs) * hd * elem;  // two planes of one head
    return (mode == KV_JOINT ? 2ull : 1ull) * (2 * rt + CT_T) * (CT_Y + 2 * rs) * (CT_X + 2 * rs) * hd * elem;
}

int attn_cell_threads(int mode) { return mode == KV_PAIR2 || mode == KV_PAIR3 ? 256 : CNT; }

// Tiles in the order rows of tiles (y), then slices, then x: the kernels
// deal consecutive tiles to one XCD, so an XCD holds bands of rows through
// all slices, and the time neighbours whose K/V rows a tile stages sit in
// its own L2 (only the rows next to a band come from another XCD).
void build_cell_tiles(const std::vector<Seg>& segs, std::vector<CellTile>& tiles) {
    tiles.clear();
    for (size_t i = 0; i < segs.size(); i++)
        for (int y = 0; y < segs[i].ny; y += CT_Y)
            for (int t = 0; t < segs[i].T; t += CT_T)
                for (int x = 0; x < segs[i].nx; x += CT_X) tiles.push_back(CellTile{(int)i, t, y, x});
}

template <class T>
static void launch(hipStream_t s, const T* qkv, T* out, int dim, int heads, int rs, int rt, const Seg* segs,
                   const CellTile* tiles, int n_tiles, int mode) {
    if (n_tiles <= 0) return;
    const int hd = dim / heads;
    const void* f = pick<T>(hd, rs, rt, mode);
    const size_t smem = attn_cell_smem(hd, rs, rt, sizeof(T), mode);
    if (!f || smem > 160 * 1024) fail(J2K_ERR_LIMIT, "attention: no LDS-tiled kernel for this head size / stencil");
    void* args[] = {(void*)&qkv, (void*)&out, (void*)&dim, (void*)&heads, (void*)&segs, (void*)&tiles};
    HIPCHK(hipLaunchKernel(f, dim3(n_tiles), dim3(attn_cell_threads(mode)), args, smem, s));
}

void attention_cell_f32(hipStream_t s, const float* qkv, float* out, int dim, int heads, int rs, int rt,
                        const Seg* segs, const CellTile* tiles, int n_tiles, int mode) {
    launch<float>(s, qkv, out, dim, heads, rs, rt, segs, tiles, n_tiles, mode);
}

void attention_cell_bf16(hipStream_t s, const uint16_t* qkv, uint16_t* out, int dim, int heads, int rs, int rt,
                         const Seg* segs, const CellTile* tiles, int n_tiles, int mode) {
    launch<uint16_t>(s, qkv, out, dim, heads, rs, rt, segs, tiles, n_tiles, mode);
}

}  // namespace j2k
