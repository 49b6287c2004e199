// The reference shallow-water solver (crates/solvers/src/shallow_water.rs)
// on the device, and dump-dataset's trajectory loop
// (crates/engine-cli/src/dump_dataset.rs:168-193) over many states at once.
//
// Arithmetic: every cell update is the reference's f32 expression, operand
// for operand and in the same order, with contraction off (the pragma below
// and -ffp-contract=off on this file), so each output is bit-identical to
// ShallowWater::step. Face fluxes are recomputed by both cells that share
// the face; the same operands give the same value, so nothing changes.
//
// Both kernels run one routine, the row march (below): a wave owns a band
// of columns and walks down its rows with the neighbours in registers.
//  - resident: one workgroup owns one state for all of its ticks in LDS
//    (updated in place, 12 B per cell, for grids up to 64 columns and
//    13,653 cells; two copies, 24 B per cell, up to 6,826 cells beyond);
//    HBM sees the state once in, once out, plus the recorded frames.
//    This is the shape of the rollout truth partner and of dump-dataset
//    (64x64 grids).
//  - tiled: one launch per step for larger grids, reading and writing HBM
//    directly: 24 B per cell and step (h, u, v in and out).
#include <algorithm>
#include <cstring>
#include <vector>

#include "abi.h"
#include "common.h"
#include "score.h"

#pragma clang fp contract(off)

namespace j2k {
namespace {

// Per-grid scalars, computed on the host as the reference computes them
// (shallow_water.rs:175, 184-185): r = dt/dx, 1/(2dx), 1/(dx*dx), in f32.
struct SwConst {
    int nx, ny;
    float r, inv2dx, invdx2, dt, g, nu;
};

// Which ticks, frames and drops a launch covers (see j2k_sw_trajectories_dev).
struct SwRun {
    int batch, max_ticks;
    const int* ticks;        // per state, or NULL: every state runs max_ticks
    const int64_t* offsets;  // frame base per state (floats) when traj != NULL
    float* traj;
    const int* drop_begin;   // [batch + 1], or NULL
    const j2k_sw_drop* drops;
    const float* fields;
};

__device__ __forceinline__ int state_ticks(const SwRun& run, int b) {
    return run.ticks ? min(run.ticks[b], run.max_ticks) : run.max_ticks;
}

// shallow_water.rs:159: the flux through one face, ((0.5*(ha+hb))*0.5)*(ua+ub).
__device__ __forceinline__ float sw_flux(float ha, float hb, float ua, float ub) {
    return 0.5f * (ha + hb) * 0.5f * (ua + ub);
}

// The drops of state b (dump_dataset.rs:85-95): the slice drops[d0, d1) of
// the schedule and whether any of them lands at tick `tick`; read once per
// wave and step, outside the row loop.
struct SwDrops {
    int d0, d1;
    bool any;
};

__device__ __forceinline__ SwDrops sw_drops_of(const SwRun& run, int b, int tick) {
    SwDrops d{0, 0, false};
    if (run.drop_begin) {
        d.d0 = run.drop_begin[b];
        d.d1 = run.drop_begin[b + 1];
        for (int i = d.d0; i < d.d1; i++) d.any |= run.drops[i].tick == tick;
    }
    return d;
}

// h after the drops of `tick`, in schedule order; cell is the cell's index.
__device__ __forceinline__ float sw_drops(const SwRun& run, const SwDrops& d, int tick, size_t n, int cell, float h) {
    if (!d.any) return h;
    for (int i = d.d0; i < d.d1; i++) {
        const j2k_sw_drop dr = run.drops[i];
        if (dr.tick == tick) h += run.fields[(size_t)dr.field * n + cell];
    }
    return h;
}

// Frame t of a state whose frames start at fb (see j2k_sw_trajectories_dev).
__device__ __forceinline__ void sw_frame(float* fb, int frames, int t, size_t n, int cell, float h, float u, float v) {
    float* f = fb + (size_t)t * n + cell;
    const size_t plane = (size_t)frames * n;
    f[0] = h;
    f[plane] = u;
    f[2 * plane] = v;
}

// ---- the row march ------------------------------------------------------------
//
// A wave owns a band of columns, one per lane, and walks down the rows
// y0 .. y1-1 of one state. Rows y-1 .. y+2 of h, u, v and h_new of rows
// y-1 .. y+1 stay in registers. A lane's x-neighbours come from the lanes
// beside it (whole-wave DPP shifts, wave_shl/wave_shr by one lane), so each
// x-face flux is computed once, by the lane west of the face, and each
// y-face flux once per column, and both cells of a face read that one value:
// the value the reference keeps in its fx / fy arrays.

// The value lane l+1 (from_east) or lane l-1 (from_west) holds; the end
// lanes of the wave receive 0 (bound_ctrl: no separate zeroing).
__device__ __forceinline__ float from_east(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_west(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x138, 0xf, 0xf, true));
}

struct SwLane {
    int gx;       // this lane's column
    int cx;       // gx clamped into the grid: the column its loads read
    bool in;      // 0 <= gx < nx
    bool ww, we;  // gx == 0 (west wall), gx == nx - 1 (east wall)
};

struct SwRow {
    float h, u, v;
};

// h_new of a row (shallow_water.rs:153-181), given the fluxes through its
// south (fyd) and north (fyu) faces; wall faces carry zero flux.
__device__ __forceinline__ float sw_hrow(const SwConst& k, const SwLane& L, const SwRow& r, float fyd, float fyu) {
    const float he = from_east(r.h), ue = from_east(r.u);
    const float fxr = L.we ? 0.0f : sw_flux(r.h, he, r.u, ue);
    const float fxw = from_west(fxr);
    const float fxl = L.ww ? 0.0f : fxw;
    const float dfx = fxr - fxl;
    const float dfy = fyu - fyd;
    return r.h - k.r * (dfx + dfy);
}

// u_new, v_new of row y (shallow_water.rs:183-230) from rows y-1, y, y+1
// (rm, rc, rp) and h_new there (hm, hc, hp). Ghosts mirror the edge cell
// (mirror(), shallow_water.rs:115-124): h even, u odd across the west and
// east walls and even across the south and north ones, v the other way.
__device__ __forceinline__ void sw_mrow(const SwConst& k, const SwLane& L, bool ys, bool yn, const SwRow& rm,
                                        const SwRow& rc, const SwRow& rp, float hm, float hc, float hp, float& un,
                                        float& vn) {
    const float hw = from_west(hc), he = from_east(hc);
    const float uw = from_west(rc.u), ue = from_east(rc.u);
    const float vw = from_west(rc.v), ve = from_east(rc.v);
    const float hl = L.ww ? hc : hw;
    const float hr = L.we ? hc : he;
    const float hd = ys ? hc : hm;
    const float hu = yn ? hc : hp;
    const float ul = L.ww ? -1.0f * rc.u : uw;
    const float ur = L.we ? -1.0f * rc.u : ue;
    const float ud = ys ? rc.u : rm.u;
    const float uu = yn ? rc.u : rp.u;
    const float vl = L.ww ? rc.v : vw;
    const float vr = L.we ? rc.v : ve;
    const float vd = ys ? -1.0f * rc.v : rm.v;
    const float vu = yn ? -1.0f * rc.v : rp.v;
    const float u0 = rc.u, v0 = rc.v;

    const float dudx = (ur - ul) * k.inv2dx;
    const float dudy = (uu - ud) * k.inv2dx;
    const float dvdx = (vr - vl) * k.inv2dx;
    const float dvdy = (vu - vd) * k.inv2dx;
    const float dhdx = (hr - hl) * k.inv2dx;
    const float dhdy = (hu - hd) * k.inv2dx;

    const float lap_u = (ul + ur + ud + uu - 4.0f * u0) * k.invdx2;
    const float lap_v = (vl + vr + vd + vu - 4.0f * v0) * k.invdx2;

    un = u0 + k.dt * (-u0 * dudx - v0 * dudy - k.g * dhdx + k.nu * lap_u);
    vn = v0 + k.dt * (-u0 * dvdx - v0 * dvdy - k.g * dhdy + k.nu * lap_v);
}

// One step of rows y0 .. y1-1 for this wave's band: fetch(y) returns this
// lane's h, u, v of row y (y past the grid's last row is clamped by the
// caller's min(), never used), emit(y, live, h_new, u_new, v_new) takes the
// results, live = y < y1 (rows past y1 must not be stored). Row y needs rows
// y-1 .. y+2 and h_new of rows y-1 .. y+1.
//
// D > 0 (HBM) keeps D more rows in flight: the rows live in a ring of
// NB = D + 3 registers sets, row r in slot (r - y0 + 1) mod NB, and the loop
// runs whole groups of NB rows, unrolled, so every slot index is static: a
// row never moves between registers and the slot of row y-1 is refilled
// with row y + NB - 1 as soon as row y is done. The row body has no branch
// (rows past y1 in the last group are computed and
// dropped by emit; the values that only matter at the grid's edge are
// selected), so each wait on a ring load lands about D rows after the load
// was issued.
// D = 0 (LDS) fetches row y+2 when row y needs it.
template <int D, class Fetch, class Emit>
__device__ __forceinline__ void sw_march(const SwConst& k, const SwLane& L, int y0, int y1, Fetch fetch, Emit emit) {
    const int last = k.ny - 1;
    // the last row this chunk reads: rows past it only feed rows past y1,
    // which are dropped, so their fetches repeat it (a cache hit, no new
    // HBM line)
    const int lim = min(y1 + 1, last);
    constexpr int NB = D + 3;
    SwRow ring[D > 0 ? NB : 1];
    if constexpr (D > 0) {
        // the rows ahead first, in row order, so they are the oldest loads
#pragma unroll
        for (int j = 3; j < NB; j++) {
            ring[j] = fetch(min(y0 + j - 1, lim));
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // h_new of row y0-1 (needs rows y0-2 .. y0) and the flux south of y0
    const SwRow r0 = fetch(y0), r1 = fetch(min(y0 + 1, lim));
    SwRow rsouth{0.0f, 0.0f, 0.0f};
    float hm = 0.0f, fys = 0.0f;
    if (y0 > 0) {
        rsouth = fetch(y0 - 1);
        fys = sw_flux(rsouth.h, r0.h, rsouth.v, r0.v);
        float fyss = 0.0f;
        if (y0 > 1) {
            const SwRow r2 = fetch(y0 - 2);
            fyss = sw_flux(r2.h, rsouth.h, r2.v, rsouth.v);
        }
        hm = sw_hrow(k, L, rsouth, fyss, fys);
    }
    float fy = y0 < last ? sw_flux(r0.h, r1.h, r0.v, r1.v) : 0.0f;  // face north of the current row
    float hc = sw_hrow(k, L, r0, fys, fy);
    // the momentum and the h_new of the next row, shared by both forms
    auto body = [&](int y, const SwRow& rm, const SwRow& rc, const SwRow& rp, const SwRow& rq) {
        const float fq = sw_flux(rp.h, rq.h, rp.v, rq.v);
        const float fyu = y + 2 <= last ? fq : 0.0f;
        const float hp = sw_hrow(k, L, rp, fy, fyu);  // past the last row: unused (the last row mirrors)
        float un, vn;
        sw_mrow(k, L, y == 0, y == last, rm, rc, rp, hm, hc, hp, un, vn);
        emit(y, y < y1, hc, un, vn);
        fy = fyu;
        hm = hc;
        hc = hp;
    };
    if constexpr (D == 0) {
        SwRow rm = rsouth, rc = r0, rp = r1;
        for (int y = y0; y < y1; y++) {
            const SwRow rq = fetch(min(y + 2, lim));
            body(y, rm, rc, rp, rq);
            rm = rc;
            rc = rp;
            rp = rq;
        }
    } else {
        ring[0] = rsouth;
        ring[1] = r0;
        ring[2] = r1;
        for (int yg = y0; yg < y1; yg += NB) {
#pragma unroll
            for (int j = 0; j < NB; j++) {
                const int y = yg + j;
                body(y, ring[j], ring[(j + 1) % NB], ring[(j + 2) % NB], ring[(j + 3) % NB]);
                ring[j] = fetch(min(y + NB - 1, lim));
            }
        }
    }
}

// How a grid splits into bands and row chunks: nx <= 64 is one band of 64
// lanes (lane = column); a wider grid takes bands of 60 output columns with
// two halo lanes on each side. A lane off the grid loads its clamped column
// (in bounds, never used: lanes on a wall read no neighbour across it).
struct SwBands {
    int bands, bw, off, rows, chunks;
};

__device__ __forceinline__ SwLane sw_lane(const SwConst& k, const SwBands& g, int band, int lane, bool& out) {
    SwLane L;
    L.gx = band * g.bw + g.off + lane;
    L.cx = min(max(L.gx, 0), k.nx - 1);
    L.in = L.gx >= 0 && L.gx < k.nx;
    L.ww = L.gx == 0;
    L.we = L.gx == k.nx - 1;
    out = L.in && lane >= -g.off && lane < 64 + g.off;
    return L;
}

constexpr int SW_LDS_BYTES = 160 * 1024;
constexpr int SW_LDS_CELLS_INPLACE = SW_LDS_BYTES / 12;  // one state, grids up to 64 columns
constexpr int SW_LDS_CELLS = SW_LDS_BYTES / 24;          // two states, wider grids

// One workgroup per state, every tick inside LDS; W waves share the
// state's (band, row chunk) items.
//  - INPLACE (one band, nx <= 64): each wave owns a chunk of rows and
//    overwrites them in place (the march has read a row before it writes
//    it), except its first and last two rows, which the neighbouring chunks
//    read as their halo: those wait in registers for a barrier, then a
//    second barrier ends the step. 12 B of LDS per cell.
//  - otherwise bands share halo columns, so the state ping-pongs between
//    two LDS copies, one barrier per step. 24 B per cell.
template <int W, bool INPLACE, bool EXTRA>
__global__ __launch_bounds__(64 * W) void k_sw_resident(SwConst k, float* gh, float* gu, float* gv, SwRun run,
                                                        SwBands g) {
    extern __shared__ float lds[];
    const int n = k.nx * k.ny;
    const int b = blockIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int T = state_ticks(run, b);
    float* fb = EXTRA && run.traj ? run.traj + run.offsets[b] : nullptr;
    float* A = lds;
    float* B = INPLACE ? lds : lds + 3 * n;
    gh += (size_t)b * n;
    gu += (size_t)b * n;
    gv += (size_t)b * n;
    for (int i = threadIdx.x; i < n; i += 64 * W) {
        const float h = gh[i], u = gu[i], v = gv[i];
        A[i] = h;
        A[n + i] = u;
        A[2 * n + i] = v;
        if (fb) sw_frame(fb, T + 1, 0, n, i, h, u, v);
    }
    __syncthreads();
    const int items = g.bands * g.chunks;
    SwRow keep[4] = {};  // in place: a chunk's boundary rows, written after the barrier
    int kept_y0 = 0, kept_y1 = 0, kept_i = 0;
    bool kept_out = false;
    for (int t = 0; t < T; t++) {
        if (INPLACE) {
            // items <= W: at most one chunk per wave. A chunk's first two and
            // last two rows are the only ones another wave reads (as its
            // halo): their new values wait in registers until every wave has
            // finished reading, the other rows are written in place at once.
            if (wave < items) {
                bool out;
                const SwLane L = sw_lane(k, g, 0, lane, out);
                const int y0 = wave * g.rows, y1 = min(y0 + g.rows, k.ny);
                const SwDrops dr = EXTRA ? sw_drops_of(run, b, t + 1) : SwDrops{0, 0, false};
                auto fetch = [&](int y) -> SwRow {
                    const int i = min(max(y, 0), k.ny - 1) * k.nx + L.cx;
                    return SwRow{A[i], A[n + i], A[2 * n + i]};
                };
                auto emit = [&](int y, bool live, float h, float u, float v) {
                    if (!live) return;
                    const int i = y * k.nx + L.cx;
                    if (EXTRA) h = sw_drops(run, dr, t + 1, n, i, h);
                    if (EXTRA && fb && out) sw_frame(fb, T + 1, t + 1, n, i, h, u, v);
                    const SwRow r{h, u, v};
                    if (y == y0) keep[0] = r;
                    else if (y == y0 + 1) keep[1] = r;
                    else if (y == y1 - 2) keep[2] = r;
                    else if (y == y1 - 1) keep[3] = r;
                    else if (out) {
                        A[i] = h;
                        A[n + i] = u;
                        A[2 * n + i] = v;
                    }
                };
                sw_march<0>(k, L, y0, y1, fetch, emit);
                kept_y0 = y0;
                kept_y1 = y1;
                kept_out = out;
                kept_i = L.cx;
            }
            __syncthreads();  // every wave has read its halo rows
            if (wave < items) {
                const int y0 = kept_y0, y1 = kept_y1;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int y = j < 2 ? y0 + j : y1 - 4 + j;
                    // rows of a chunk under four rows are kept once, under
                    // their first slot
                    const bool mine = y < y1 && (j < 2 || y >= y0 + 2);
                    if (mine && kept_out) {
                        const int i = y * k.nx + kept_i;
                        A[i] = keep[j].h;
                        A[n + i] = keep[j].u;
                        A[2 * n + i] = keep[j].v;
                    }
                }
            }
            __syncthreads();
        } else {
            const SwDrops dr = EXTRA ? sw_drops_of(run, b, t + 1) : SwDrops{0, 0, false};
            for (int it = wave; it < items; it += W) {
                const int chunk = it / g.bands, band = it - chunk * g.bands;
                bool out;
                const SwLane L = sw_lane(k, g, band, lane, out);
                const float* src = A;
                float* dst = B;
                auto fetch = [&](int y) -> SwRow {
                    const int i = y * k.nx + L.cx;
                    return SwRow{src[i], src[n + i], src[2 * n + i]};
                };
                auto emit = [&](int y, bool live, float h, float u, float v) {
                    if (!live || !out) return;
                    const int i = y * k.nx + L.gx;
                    h = sw_drops(run, dr, t + 1, n, i, h);
                    dst[i] = h;
                    dst[n + i] = u;
                    dst[2 * n + i] = v;
                    if (fb) sw_frame(fb, T + 1, t + 1, n, i, h, u, v);
                };
                const int y0 = chunk * g.rows;
                sw_march<0>(k, L, y0, min(y0 + g.rows, k.ny), fetch, emit);
            }
            __syncthreads();
            float* tmp = A;
            A = B;
            B = tmp;
        }
    }
    for (int i = threadIdx.x; i < n; i += 64 * W) {
        gh[i] = A[i];
        gu[i] = A[n + i];
        gv[i] = A[2 * n + i];
    }
}

struct SwIO {
    const float *h, *u, *v;
};
struct SwOut {
    float *h, *u, *v;
};

constexpr int SW_PREFETCH = 4;  // rows in flight ahead of the march (HBM)
// rows a wave of `rows` output rows may fetch: two above, the padded last
// group and the ring's refills below
__host__ __device__ constexpr int sw_span(int rows) { return rows + 2 * (SW_PREFETCH + 3) + 2; }

// Raw buffer access (MUBUF): a resource over `bytes` bytes from base, a lane
// byte offset and a wave-uniform one. A lane whose offset lies past the
// resource's range reads 0 and its store is dropped by the hardware's range
// check, which is how lanes that own no output skip their stores without a
// branch (a branch around stores makes the compiler assume the fewest
// memory operations in flight and wait early for the row loads).
constexpr unsigned SW_RSRC_WORD3 = 0x00020000;  // gfx9 raw buffer, 32-bit elements
constexpr int SW_OFF_NONE = 0x40000000;         // a lane offset past any resource here

__device__ __forceinline__ __amdgpu_buffer_rsrc_t sw_rsrc(const float* base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), 0, bytes, SW_RSRC_WORD3);
}
__device__ __forceinline__ float sw_bload(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void sw_bstore(__amdgpu_buffer_rsrc_t r, int voff, int soff, float x) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, x), r, voff, soff, 0);
}

// Step t (0-based) of every state whose ticks exceed t, src -> dst in HBM.
// Each wave marches one (state, row chunk, band) item; items run state by
// state, chunk by chunk, so a chunk's halo rows were read by the chunk
// before it on the same XCD (blocks are dealt to the XCDs in contiguous
// runs). A state already past its ticks is carried over unchanged, which
// keeps the ping-pong buffers in step. Rows are read and written through
// buffer resources over the wave's rows (at most sw_span(rows) rows, under
// 2 GiB: checked on the host).
template <bool EXTRA>
__global__ __launch_bounds__(256) void k_sw_band(SwConst k, SwIO src, SwOut dst, int t, SwRun run, SwBands g) {
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const int item = __builtin_amdgcn_readfirstlane(wg * 4 + (threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int per_state = g.chunks * g.bands;
    if (item >= per_state * run.batch) return;
    const int b = item / per_state;
    const int r = item - b * per_state;
    const int chunk = r / g.bands, band = r - chunk * g.bands;
    bool out;
    const SwLane L = sw_lane(k, g, band, lane, out);
    const size_t n = (size_t)k.nx * k.ny, base = (size_t)b * n;
    const int y0 = chunk * g.rows, y1 = min(y0 + g.rows, k.ny);
    const int row_bytes = k.nx * 4;
    // source rows [ylo, yhi): every row the march fetches
    const int ylo = max(y0 - 2, 0), yhi = min(k.ny, y0 + sw_span(g.rows) - 2);
    const int sbytes = (yhi - ylo) * row_bytes, dbytes = (y1 - y0) * row_bytes;
    const size_t so = base + (size_t)ylo * k.nx, dso = base + (size_t)y0 * k.nx;
    const auto rh = sw_rsrc(src.h + so, sbytes), ru = sw_rsrc(src.u + so, sbytes), rv = sw_rsrc(src.v + so, sbytes);
    const auto wh = sw_rsrc(dst.h + dso, dbytes), wu = sw_rsrc(dst.u + dso, dbytes), wv = sw_rsrc(dst.v + dso, dbytes);
    const int lin = L.cx * 4, lout = out ? L.gx * 4 : SW_OFF_NONE;
    auto fetch = [&](int y) -> SwRow {
        const int so_y = (y - ylo) * row_bytes;
        return SwRow{sw_bload(rh, lin, so_y), sw_bload(ru, lin, so_y), sw_bload(rv, lin, so_y)};
    };
    const int T = state_ticks(run, b);
    if (t >= T) {
        for (int y = y0; y < y1; y++) {
            const SwRow row = fetch(y);
            const int so_y = (y - y0) * row_bytes;
            sw_bstore(wh, lout, so_y, row.h);
            sw_bstore(wu, lout, so_y, row.u);
            sw_bstore(wv, lout, so_y, row.v);
        }
        return;
    }
    const SwDrops dr = EXTRA ? sw_drops_of(run, b, t + 1) : SwDrops{0, 0, false};
    float* fb = EXTRA && run.traj ? run.traj + run.offsets[b] : nullptr;
    auto emit = [&](int y, bool live, float h, float u, float v) {
        const int cell = min(y, k.ny - 1) * k.nx + L.cx;
        if (EXTRA) h = sw_drops(run, dr, t + 1, n, cell, h);
        // a row past y1 goes to no cell: every lane's offset is out of range
        const int lo = live ? lout : SW_OFF_NONE;
        const int so_y = live ? (y - y0) * row_bytes : 0;
        sw_bstore(wh, lo, so_y, h);
        sw_bstore(wu, lo, so_y, u);
        sw_bstore(wv, lo, so_y, v);
        if (EXTRA && fb && live && out) sw_frame(fb, T + 1, t + 1, n, cell, h, u, v);
    };
    sw_march<SW_PREFETCH>(k, L, y0, y1, fetch, emit);
}

// Frame 0 of every state (its input) for the HBM path.
__global__ __launch_bounds__(256) void k_sw_frame0(const float* h, const float* u, const float* v, size_t n, int chunks,
                                                   SwRun run) {
    const int b = blockIdx.x / chunks, c = blockIdx.x - b * chunks;
    const int T = state_ticks(run, b);
    const size_t base = (size_t)b * n;
    for (size_t i = (size_t)c * 256 + threadIdx.x; i < n; i += (size_t)chunks * 256)
        sw_frame(run.traj + run.offsets[b], T + 1, 0, n, (int)i, h[base + i], u[base + i], v[base + i]);
}

// ---- host side ------------------------------------------------------------------

struct Solver {
    j2k_sw_params p{};
    SwConst k{};
    int device = 0;
    int kernel = 0;  // 0 auto, 1 resident, 2 tiled
    int cus = 256;   // compute units of the device
    hipStream_t stream = nullptr;
    // device staging: ping-pong states for the tiled path, host-call copies
    float* pp = nullptr;
    size_t pp_floats = 0;
    float* io = nullptr;
    size_t io_floats = 0;
    // frames of the model loops' truth partner, and its one frame offset (0)
    float* fr = nullptr;
    size_t fr_floats = 0;
    int64_t* off0 = nullptr;

    ~Solver() {
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        if (pp) (void)hipFree(pp);
        if (io) (void)hipFree(io);
        if (fr) (void)hipFree(fr);
        if (off0) (void)hipFree(off0);
        if (stream) (void)hipStreamDestroy(stream);
        if (prev >= 0) (void)hipSetDevice(prev);
    }
    size_t cells() const { return (size_t)p.nx * p.ny; }
};

struct Dev {
    int prev = -1;
    explicit Dev(int d) {
        HIPCHK(hipGetDevice(&prev));
        if (prev != d) HIPCHK(hipSetDevice(d));
    }
    ~Dev() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

Solver& solver_of(void* h) {
    if (!h) fail(J2K_ERR_INVALID, "solver handle is NULL");
    return *static_cast<Solver*>(h);
}

void need(const void* p, const char* what) {
    if (!p) fail(J2K_ERR_INVALID, std::string(what) + " is NULL");
}

void check_params(const j2k_sw_params* p) {
    need(p, "params");
    if (p->nx < 1 || p->ny < 1) fail(J2K_ERR_INVALID, "solver grid needs nx >= 1 and ny >= 1");
    if ((long long)p->nx * p->ny >= (1LL << 31)) fail(J2K_ERR_INVALID, "solver grid over 2^31 cells");
}

template <class T>
T* grow(T*& buf, size_t& have, size_t need_count) {
    if (need_count > have) {
        if (buf) HIPCHK(hipFree(buf));
        buf = nullptr;
        have = 0;
        HIPCHK(hipMalloc(&buf, need_count * sizeof(T)));
        have = need_count;
    }
    return buf;
}

// Bands of a grid (see SwBands) with row chunks of `rows`.
SwBands bands_of(const SwConst& k, int rows) {
    SwBands g{};
    if (k.nx <= 64) {
        g.bands = 1, g.bw = 64, g.off = 0;
    } else {
        g.bands = cdiv(k.nx, 60), g.bw = 60, g.off = -2;
    }
    g.rows = std::max(1, std::min(rows, k.ny));
    g.chunks = cdiv(k.ny, g.rows);
    return g;
}

bool resident_fits(const j2k_sw_params& p) {
    const long long n = (long long)p.nx * p.ny;
    return p.nx <= 64 ? n <= SW_LDS_CELLS_INPLACE : n <= SW_LDS_CELLS && cdiv(p.nx, 60) <= 16;
}

bool use_resident(const Solver& s) {
    const bool fits = resident_fits(s.p);
    if (s.kernel == 1 && !fits)
        fail(J2K_ERR_LIMIT, "resident solver kernel: a " + std::to_string(s.p.nx) + "x" + std::to_string(s.p.ny) +
                                " grid does not fit one workgroup's LDS (up to " +
                                std::to_string(SW_LDS_CELLS_INPLACE) + " cells at 64 columns or fewer, " +
                                std::to_string(SW_LDS_CELLS) + " beyond)");
    return s.kernel == 1 || (s.kernel == 0 && fits);
}

// Waves per state: a batch that fills every CU several times over runs 8,
// a smaller one 16 (short row chunks, the lowest time per tick); rows split so
// that the (band, chunk) items about match the waves (in place: at most one
// chunk per wave).
void run_resident(Solver& s, hipStream_t st, float* h, float* u, float* v, const SwRun& run) {
    const int W = run.batch >= 2 * s.cus ? 8 : 16;
    const bool inplace = s.p.nx <= 64;
    const int per_band = std::max(1, W / (inplace ? 1 : cdiv(s.p.nx, 60)));
    const SwBands g = bands_of(s.k, cdiv(s.p.ny, per_band));
    const bool extra = run.traj || run.drop_begin;
    using Fn = void (*)(SwConst, float*, float*, float*, SwRun, SwBands);
    static const Fn fns[2][2][2] = {
        {{k_sw_resident<8, false, false>, k_sw_resident<8, false, true>},
         {k_sw_resident<8, true, false>, k_sw_resident<8, true, true>}},
        {{k_sw_resident<16, false, false>, k_sw_resident<16, false, true>},
         {k_sw_resident<16, true, false>, k_sw_resident<16, true, true>}}};
    const Fn fn = fns[W == 16][inplace][extra];
    const size_t smem = (inplace ? 12 : 24) * s.cells();
    if (smem > 64 * 1024) HIPCHK(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    hipLaunchKernelGGL(fn, dim3(run.batch), dim3(64 * W), smem, st, s.k, h, u, v, run, g);
    LAUNCHCHK("k_sw_resident");
}

// One HBM launch per step: user -> A -> B -> A ... -> user, the last step
// writing the caller's buffers (a single step first copies them to A).
void run_tiled(Solver& s, hipStream_t st, float* h, float* u, float* v, const SwRun& run) {
    const size_t n = s.cells();
    const size_t state = (size_t)run.batch * n;
    if (run.traj) {
        const int chunks = (int)std::min<size_t>(cdiv((long)n, 256), 64);
        hipLaunchKernelGGL(k_sw_frame0, dim3((unsigned)(chunks * run.batch)), dim3(256), 0, st, h, u, v, n, chunks, run);
        LAUNCHCHK("k_sw_frame0");
    }
    if (run.max_ticks == 0) return;
    float* pp = grow(s.pp, s.pp_floats, 6 * state);
    SwOut A{pp, pp + state, pp + 2 * state}, B{pp + 3 * state, pp + 4 * state, pp + 5 * state}, U{h, u, v};
    SwIO src{h, u, v};
    if (run.max_ticks == 1) {
        HIPCHK(hipMemcpyAsync(A.h, h, state * 4, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(A.u, u, state * 4, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(A.v, v, state * 4, hipMemcpyDeviceToDevice, st));
        src = SwIO{A.h, A.u, A.v};
    }
    // rows per wave: enough waves to fill about 3/4 of the GPU's wave slots
    // (8 per SIMD at this kernel's register count) in one round, at least 16
    const long long slots = 6LL * 4 * s.cus;
    const long long bands = s.p.nx <= 64 ? 1 : cdiv(s.p.nx, 60);
    const long long rows = std::max<long long>(16, (bands * s.p.ny * run.batch + slots - 1) / slots);
    const SwBands g = bands_of(s.k, (int)std::min<long long>(rows, s.p.ny));
    if ((long long)sw_span(g.rows) * s.p.nx * 4 >= (1LL << 31))
        fail(J2K_ERR_LIMIT, "solver grid too large for the HBM kernel's row buffers (" + std::to_string(s.p.nx) +
                                " columns)");
    const long long waves = (long long)g.bands * g.chunks * run.batch;
    if (waves >= (1LL << 31)) fail(J2K_ERR_LIMIT, "solver launch over 2^31 waves");
    const unsigned blocks = (unsigned)((waves + 3) / 4);
    for (int t = 0; t < run.max_ticks; t++) {
        const SwOut dst = t == run.max_ticks - 1 ? U : (t % 2 == 0 ? A : B);
        if (run.traj || run.drop_begin)
            hipLaunchKernelGGL(k_sw_band<true>, dim3(blocks), dim3(256), 0, st, s.k, src, dst, t, run, g);
        else
            hipLaunchKernelGGL(k_sw_band<false>, dim3(blocks), dim3(256), 0, st, s.k, src, dst, t, run, g);
        LAUNCHCHK("k_sw_band");
        src = SwIO{dst.h, dst.u, dst.v};
    }
}

void run(Solver& s, hipStream_t st, float* h, float* u, float* v, const SwRun& r) {
    if (r.batch == 0) return;
    if (use_resident(s)) run_resident(s, st, h, u, v, r);
    else run_tiled(s, st, h, u, v, r);
}

SwRun plain_run(int batch, int steps) {
    SwRun r{};
    r.batch = batch;
    r.max_ticks = steps;
    return r;
}

void check_batch(int batch, int steps) {
    if (batch < 0) fail(J2K_ERR_INVALID, "batch must be >= 0");
    if (steps < 0) fail(J2K_ERR_INVALID, "steps must be >= 0");
}

}  // namespace

float* sw_frames_on(void* solver, hipStream_t st, float* h, float* u, float* v, int steps) {
    Solver& s = solver_of(solver);
    Dev dg(s.device);
    if (!s.off0) {
        HIPCHK(hipMalloc(&s.off0, sizeof(int64_t)));
        const int64_t zero = 0;
        HIPCHK(hipMemcpy(s.off0, &zero, sizeof(zero), hipMemcpyHostToDevice));
    }
    float* fr = grow(s.fr, s.fr_floats, 3 * (size_t)(steps + 1) * s.cells());
    SwRun r = plain_run(1, steps);
    r.offsets = s.off0;
    r.traj = fr;
    run(s, st, h, u, v, r);
    return fr;
}

void sw_grid_of(const void* solver, int* nx, int* ny, float* depth, int* device) {
    const Solver& s = solver_of(const_cast<void*>(solver));
    *nx = s.p.nx;
    *ny = s.p.ny;
    *depth = s.p.depth;
    *device = s.device;
}

}  // namespace j2k

using namespace j2k;

extern "C" {

int j2k_sw_create(const j2k_sw_params* params, int device, void** out) {
    return guard([&] {
        check_params(params);
        need(out, "out_solver");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) fail(J2K_ERR_NO_ADAPTER, "no HIP device is visible");
        if (device < 0 || device >= count) fail(J2K_ERR_NO_DEVICE, "device index " + std::to_string(device) + " out of range");
        Dev dg(device);
        auto* s = new Solver();
        s->p = *params;
        s->device = device;
        const j2k_sw_params& p = s->p;
        s->k = SwConst{p.nx, p.ny, p.dt / p.dx, 1.0f / (2.0f * p.dx), 1.0f / (p.dx * p.dx), p.dt, p.g, p.nu};
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            s->cus = prop.multiProcessorCount;
        if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
            delete s;
            fail(J2K_ERR_NO_DEVICE, "hipStreamCreate failed");
        }
        *out = s;
    });
}

int j2k_sw_destroy(void* h) {
    return guard([&] { delete static_cast<Solver*>(h); });
}

int j2k_sw_set_kernel(void* h, int kernel) {
    return guard([&] {
        Solver& s = solver_of(h);
        if (kernel < 0 || kernel > 2) fail(J2K_ERR_INVALID, "kernel must be 0 (auto), 1 (resident) or 2 (tiled)");
        s.kernel = kernel;
    });
}

int j2k_sw_steps_dev(void* h, float* d_h, float* d_u, float* d_v, int batch, int steps, void* stream) {
    return guard([&] {
        Solver& s = solver_of(h);
        check_batch(batch, steps);
        if (batch == 0 || steps == 0) return;
        need(d_h, "h");
        need(d_u, "u");
        need(d_v, "v");
        Dev dg(s.device);
        run(s, stream ? static_cast<hipStream_t>(stream) : s.stream, d_h, d_u, d_v, plain_run(batch, steps));
    });
}

int j2k_sw_steps(void* h, float* hh, float* hu, float* hv, int batch, int steps) {
    return guard([&] {
        Solver& s = solver_of(h);
        check_batch(batch, steps);
        if (batch == 0 || steps == 0) return;
        need(hh, "h");
        need(hu, "u");
        need(hv, "v");
        Dev dg(s.device);
        const size_t bytes = (size_t)batch * s.cells() * 4;
        char* io = reinterpret_cast<char*>(grow(s.io, s.io_floats, 3 * bytes / 4));
        float* dh = reinterpret_cast<float*>(io);
        float* du = reinterpret_cast<float*>(io + bytes);
        float* dv = reinterpret_cast<float*>(io + 2 * bytes);
        hipStream_t st = s.stream;
        HIPCHK(hipMemcpyAsync(dh, hh, bytes, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(du, hu, bytes, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dv, hv, bytes, hipMemcpyHostToDevice, st));
        run(s, st, dh, du, dv, plain_run(batch, steps));
        HIPCHK(hipMemcpyAsync(hh, dh, bytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(hu, du, bytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(hv, dv, bytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    });
}

int j2k_sw_trajectories_dev(void* h, float* d_h, float* d_u, float* d_v, int batch, const int* d_ticks, int max_ticks,
                            const int64_t* d_offsets, float* d_traj, const int* d_drop_begin, const j2k_sw_drop* d_drops,
                            const float* d_fields, void* stream) {
    return guard([&] {
        Solver& s = solver_of(h);
        check_batch(batch, max_ticks);
        if (batch == 0) return;
        need(d_h, "h");
        need(d_u, "u");
        need(d_v, "v");
        if (d_traj) need(d_offsets, "offsets");
        if (d_drop_begin) {
            need(d_drops, "drops");
            need(d_fields, "fields");
        }
        Dev dg(s.device);
        SwRun r{batch, max_ticks, d_ticks, d_offsets, d_traj, d_drop_begin, d_drops, d_fields};
        run(s, stream ? static_cast<hipStream_t>(stream) : s.stream, d_h, d_u, d_v, r);
    });
}

int j2k_sw_trajectories(void* h, float* hh, float* hu, float* hv, int batch, const int* ticks, float* traj,
                        int n_drops, const int* dstf, const float* fields, int n_fields) {
    return guard([&] {
        Solver& s = solver_of(h);
        if (batch < 0) fail(J2K_ERR_INVALID, "batch must be >= 0");
        if (batch == 0) return;
        need(hh, "h");
        need(hu, "u");
        need(hv, "v");
        need(ticks, "ticks");
        if (n_drops < 0 || n_fields < 0) fail(J2K_ERR_INVALID, "n_drops and n_fields must be >= 0");
        if (n_drops > 0) {
            need(dstf, "drops");
            need(fields, "fields");
        }
        const size_t n = s.cells();
        std::vector<int64_t> off(batch);
        int64_t total = 0;
        int max_ticks = 0;
        for (int b = 0; b < batch; b++) {
            if (ticks[b] < 0) fail(J2K_ERR_INVALID, "ticks must be >= 0");
            off[b] = total;
            total += 3 * (int64_t)(ticks[b] + 1) * (int64_t)n;
            max_ticks = std::max(max_ticks, ticks[b]);
        }
        // drops grouped by state, each state's in the caller's order
        std::vector<int> begin(batch + 1, 0);
        for (int d = 0; d < n_drops; d++) {
            const int st = dstf[3 * d], tk = dstf[3 * d + 1], fi = dstf[3 * d + 2];
            if (st < 0 || st >= batch) fail(J2K_ERR_INVALID, "drop state out of range");
            if (fi < 0 || fi >= n_fields) fail(J2K_ERR_INVALID, "drop field out of range");
            (void)tk;
            begin[st + 1]++;
        }
        for (int b = 0; b < batch; b++) begin[b + 1] += begin[b];
        std::vector<j2k_sw_drop> drops(std::max(n_drops, 1));
        {
            std::vector<int> fillp(begin.begin(), begin.end() - 1);
            for (int d = 0; d < n_drops; d++) drops[fillp[dstf[3 * d]]++] = j2k_sw_drop{dstf[3 * d + 1], dstf[3 * d + 2]};
        }
        Dev dg(s.device);
        hipStream_t st = s.stream;
        const size_t sbytes = (size_t)batch * n * 4;
        const size_t fbytes = (size_t)n_fields * n * 4;
        // one device block: h u v | traj | fields | ticks | offsets | begin | drops
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        const size_t o_traj = 3 * al(sbytes), o_fields = o_traj + al((size_t)total * 4), o_ticks = o_fields + al(fbytes),
                     o_off = o_ticks + al((size_t)batch * 4), o_begin = o_off + al((size_t)batch * 8),
                     o_drops = o_begin + al((size_t)(batch + 1) * 4), end = o_drops + al(drops.size() * sizeof(j2k_sw_drop));
        char* io = reinterpret_cast<char*>(grow(s.io, s.io_floats, (end + 3) / 4));
        float* dh = reinterpret_cast<float*>(io);
        float* du = reinterpret_cast<float*>(io + al(sbytes));
        float* dv = reinterpret_cast<float*>(io + 2 * al(sbytes));
        float* dtraj = traj ? reinterpret_cast<float*>(io + o_traj) : nullptr;
        HIPCHK(hipMemcpyAsync(dh, hh, sbytes, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(du, hu, sbytes, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dv, hv, sbytes, hipMemcpyHostToDevice, st));
        if (fbytes) HIPCHK(hipMemcpyAsync(io + o_fields, fields, fbytes, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(io + o_ticks, ticks, (size_t)batch * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(io + o_off, off.data(), (size_t)batch * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(io + o_begin, begin.data(), (size_t)(batch + 1) * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(io + o_drops, drops.data(), drops.size() * sizeof(j2k_sw_drop), hipMemcpyHostToDevice, st));
        SwRun r{batch,
                max_ticks,
                reinterpret_cast<const int*>(io + o_ticks),
                reinterpret_cast<const int64_t*>(io + o_off),
                dtraj,
                n_drops ? reinterpret_cast<const int*>(io + o_begin) : nullptr,
                reinterpret_cast<const j2k_sw_drop*>(io + o_drops),
                reinterpret_cast<const float*>(io + o_fields)};
        run(s, st, dh, du, dv, r);
        HIPCHK(hipMemcpyAsync(hh, dh, sbytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(hu, du, sbytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(hv, dv, sbytes, hipMemcpyDeviceToHost, st));
        if (traj) HIPCHK(hipMemcpyAsync(traj, dtraj, (size_t)total * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    });
}

int j2k_sw_drop_field(const j2k_sw_params* p, float cx, float cy, float amp, float sigma, float* out) {
    return guard([&] {
        check_params(p);
        need(out, "out");
        const float inv = 1.0f / (2.0f * sigma * sigma);
        for (int y = 0; y < p->ny; y++)
            for (int x = 0; x < p->nx; x++) {
                const float dx = (float)x - cx;
                const float dy = (float)y - cy;
                out[(size_t)y * p->nx + x] = amp * std::exp(-(dx * dx + dy * dy) * inv);
            }
    });
}

}  // extern "C"
