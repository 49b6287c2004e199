// Mass and energy projection on the device with f64 sums
// (crates/engine/src/project.rs:9-54; eval.rs:144-156 for the order and the
// violation metric), and the cell_st per-slice variant.
//
// One 256-thread block owns one grid (or one slice): thread i visits cells
// i, i+256, ... in order, then a fixed shuffle tree and a fixed cross-wave
// order reduce the partials. The result depends only on the slice's own
// data, never on where it sits in a packed batch.
#include "common.h"
#include "kernels.h"

namespace j2k {

constexpr int PT = 256;
// Cell kernels work one block per slice (a few per request): a whole CU's
// worth of threads each keeps their dependent passes short.
constexpr int CPT = 1024;

// Block-wide f64 sum of up to 5 values; every thread gets the totals.
template <int NV, int T = PT>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*sh)[T / 64]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; j++) {
        double t = wave_sum_d(v[j]);
        if (lane == 0) sh[j][wave] = t;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NV; j++) {
        double t = 0;
#pragma unroll
        for (int w = 0; w < T / 64; w++) t += sh[j][w];
        v[j] = t;
    }
    __syncthreads();
}

struct Phys {
    float g, mean[3], stdv[3], depth, r_max;
};

// ---- patch2d grids [B, 3, n] ------------------------------------------------

__global__ __launch_bounds__(PT) void k_project_p(int n, float* __restrict__ xall, const double* __restrict__ mass_t,
                                                  int do_mass, const double* __restrict__ eprev, int do_energy,
                                                  double* __restrict__ eout, Phys ph) {
    __shared__ double sh[2][PT / 64];
    const int b = blockIdx.x;
    float* x = xall + (long)b * 3 * n;
    if (do_mass) {
        double v[1] = {0};
        for (int i = threadIdx.x; i < n; i += PT) v[0] += (double)x[i];
        block_sum<1>(v, sh);
        // project.rs:33-38: delta = target - mean, added to every h cell.
        const float delta = (float)(mass_t[b] - v[0] / (double)n);
        for (int i = threadIdx.x; i < n; i += PT) x[i] += delta;
        __syncthreads();
    }
    const float hm = ph.mean[0] + ph.depth;
    if (do_energy) {
        double v[2] = {0, 0};
        for (int i = threadIdx.x; i < n; i += PT) {
            const float h = x[i] * ph.stdv[0] + hm, u = x[n + i] * ph.stdv[1] + ph.mean[1],
                        w = x[2 * n + i] * ph.stdv[2] + ph.mean[2];
            v[0] += 0.5 * (double)ph.g * (double)h * (double)h;
            v[1] += 0.5 * (double)h * ((double)u * (double)u + (double)w * (double)w);
        }
        block_sum<2>(v, sh);
        // project.rs:43-54: scale (u, v) down to the bound, never up, K floor 1e-12.
        const double bound = eprev[b] * (double)ph.r_max;
        const double budget = fmax(bound - v[0], 0.0);
        double sc = sqrt(budget / fmax(v[1], 1e-12));
        sc = fmin(fmax(sc, 0.0), 1.0);
        const float s = (float)sc;
        for (int i = threadIdx.x; i < n; i += PT) {
            const float u = x[n + i] * ph.stdv[1] + ph.mean[1], w = x[2 * n + i] * ph.stdv[2] + ph.mean[2];
            x[n + i] = (u * s - ph.mean[1]) / ph.stdv[1];
            x[2 * n + i] = (w * s - ph.mean[2]) / ph.stdv[2];
        }
        __syncthreads();
    }
    if (eout) {
        double v[1] = {0};
        for (int i = threadIdx.x; i < n; i += PT) {
            const float h = x[i] * ph.stdv[0] + hm, u = x[n + i] * ph.stdv[1] + ph.mean[1],
                        w = x[2 * n + i] * ph.stdv[2] + ph.mean[2];
            v[0] += 0.5 * (double)ph.g * (double)h * (double)h +
                    0.5 * (double)h * ((double)u * (double)u + (double)w * (double)w);
        }
        block_sum<1>(v, sh);
        if (threadIdx.x == 0) eout[b] = v[0];
    }
}

static Phys phys_of(float g, const float* mean, const float* stdv, float depth, float r_max) {
    Phys p;
    p.g = g;
    for (int i = 0; i < 3; i++) {
        p.mean[i] = mean[i];
        p.stdv[i] = stdv[i];
    }
    p.depth = depth;
    p.r_max = r_max;
    return p;
}

void project_p(hipStream_t s, int B, int n, float* x, const double* mass_t, int do_mass, const double* eprev,
               int do_energy, double* eout, float g, const float* mean, const float* stdv, float depth, float r_max) {
    if (B <= 0) return;
    hipLaunchKernelGGL(k_project_p, dim3(B), dim3(PT), 0, s, n, x, mass_t, do_mass, eprev, do_energy, eout,
                       phys_of(g, mean, stdv, depth, r_max));
    LAUNCHCHK("project_p");
}

__global__ __launch_bounds__(PT) void k_grid_sums(int n, const float* __restrict__ xall, double* __restrict__ mass,
                                                  double* __restrict__ energy, Phys ph) {
    __shared__ double sh[2][PT / 64];
    const float* x = xall + (long)blockIdx.x * 3 * n;
    const float hm = ph.mean[0] + ph.depth;
    double v[2] = {0, 0};
    for (int i = threadIdx.x; i < n; i += PT) {
        v[0] += (double)x[i];
        const float h = x[i] * ph.stdv[0] + hm, u = x[n + i] * ph.stdv[1] + ph.mean[1],
                    w = x[2 * n + i] * ph.stdv[2] + ph.mean[2];
        v[1] += 0.5 * (double)ph.g * (double)h * (double)h +
                0.5 * (double)h * ((double)u * (double)u + (double)w * (double)w);
    }
    block_sum<2>(v, sh);
    if (threadIdx.x == 0) {
        if (mass) mass[blockIdx.x] = v[0] / (double)n;
        if (energy) energy[blockIdx.x] = v[1];
    }
}

void grid_sums(hipStream_t s, int B, int n, const float* x, double* mass, double* energy, float g, const float* mean,
               const float* stdv, float depth) {
    if (B <= 0) return;
    hipLaunchKernelGGL(k_grid_sums, dim3(B), dim3(PT), 0, s, n, x, mass, energy, phys_of(g, mean, stdv, depth, 1.0f));
    LAUNCHCHK("grid_sums");
}

// ---- cell_st slices (token-major float4 h, u, v, known) -----------------------

__device__ __forceinline__ void phys_huv(const float4& c, const Phys& ph, double& h, double& u, double& v) {
    h = (double)c.x * ph.stdv[0] + ((double)ph.mean[0] + (double)ph.depth);
    u = (double)c.y * ph.stdv[1] + ph.mean[1];
    v = (double)c.z * ph.stdv[2] + ph.mean[2];
}

__global__ __launch_bounds__(CPT) void k_cell_targets(const Slice* __restrict__ sl, const float4* __restrict__ x0,
                                                     double* __restrict__ mass_t, double* __restrict__ eref_t, Phys ph) {
    __shared__ double sh[2][CPT / 64];
    const Slice S = sl[blockIdx.x];
    double v[2] = {0, 0};
    for (int i = threadIdx.x; i < S.n; i += CPT) {
        const float4 c = x0[S.start + i];
        double h, u, w;
        phys_huv(c, ph, h, u, w);
        v[0] += (double)c.x;
        v[1] += 0.5 * (double)ph.g * h * h + 0.5 * h * (u * u + w * w);
    }
    block_sum<2, CPT>(v, sh);
    if (threadIdx.x == 0) {
        mass_t[blockIdx.x] = v[0] / (double)S.n;
        eref_t[blockIdx.x] = v[1];
    }
}

// One request's pack and targets in one launch: block b packs slice b's
// tokens (grid planes [3][M] and the known mask in token order) into x0 and
// x1 and sums them as k_cell_targets does, in its order. With ml.meta, one
// more block (the last) makes the request's meta vector as k_meta_linear.
struct MetaLin {
    int md = 0, dim = 0;
    const float *meta = nullptr, *w = nullptr, *b = nullptr;
    float* out = nullptr;
};
__global__ __launch_bounds__(CPT) void k_cell_pack_targets(const Slice* __restrict__ sl, int M,
                                                          const float* __restrict__ grid,
                                                          const uint8_t* __restrict__ known, float4* __restrict__ x0,
                                                          float4* __restrict__ x1, double* __restrict__ mass_t,
                                                          double* __restrict__ eref_t, Phys ph, MetaLin ml) {
    __shared__ double sh[2][CPT / 64];
    if (ml.meta && blockIdx.x == gridDim.x - 1) {
        for (int o = threadIdx.x; o < ml.dim; o += CPT) ml.out[o] = meta_out(ml.md, ml.dim, ml.meta, ml.w, ml.b, o);
        return;
    }
    const Slice S = sl[blockIdx.x];
    double v[2] = {0, 0};
    for (int i = threadIdx.x; i < S.n; i += CPT) {
        const long m = (long)S.start + i;
        const float4 c = make_float4(grid[m], grid[M + m], grid[2L * M + m], known[m] ? 1.0f : 0.0f);
        x0[m] = c;
        x1[m] = c;
        double h, u, w;
        phys_huv(c, ph, h, u, w);
        v[0] += (double)c.x;
        v[1] += 0.5 * (double)ph.g * h * h + 0.5 * h * (u * u + w * w);
    }
    block_sum<2, CPT>(v, sh);
    if (threadIdx.x == 0) {
        mass_t[blockIdx.x] = v[0] / (double)S.n;
        eref_t[blockIdx.x] = v[1];
    }
}

void cell_pack_targets(hipStream_t s, int n_slices, const Slice* sl, int M, const float* grid, const uint8_t* known,
                       float* x0, float* x1, double* mass_t, double* eref_t, float g, const float* mean,
                       const float* stdv, float depth, int meta_dim, int dim, const float* meta, const float* meta_w,
                       const float* meta_b, float* mvec) {
    MetaLin ml;
    if (meta) ml = MetaLin{meta_dim, dim, meta, meta_w, meta_b, mvec};
    if (n_slices <= 0 && !meta) return;
    hipLaunchKernelGGL(k_cell_pack_targets, dim3(std::max(n_slices, 0) + (meta ? 1 : 0)), dim3(CPT), 0, s, sl, M, grid,
                       known, (float4*)x0, (float4*)x1, mass_t, eref_t, phys_of(g, mean, stdv, depth, 1.0f), ml);
    LAUNCHCHK("cell_pack_targets");
}

void cell_targets(hipStream_t s, int n_slices, const Slice* sl, const float* x0, double* mass_t, double* eref_t,
                  float g, const float* mean, const float* stdv, float depth) {
    if (n_slices <= 0) return;
    hipLaunchKernelGGL(k_cell_targets, dim3(n_slices), dim3(CPT), 0, s, sl, (const float4*)x0, mass_t, eref_t,
                       phys_of(g, mean, stdv, depth, 1.0f));
    LAUNCHCHK("cell_targets");
}

// k_cell_project for a slice of at most CPT*K cells: the same passes in the
// same order, with each thread's cells (tid, tid + CPT, ...) held in
// registers from the first load to the one store, so a pass costs its
// arithmetic and one block reduction instead of a round trip to memory.
template <int K>
__device__ __forceinline__ void cell_project_regs(const Slice& S, float4* __restrict__ x,
                                                  const float4* __restrict__ p, double mass, double eref,
                                                  unsigned* __restrict__ resid, unsigned long long* __restrict__ viol,
                                                  const Phys& ph, double (*sh)[CPT / 64], float* shm,
                                                  float* __restrict__ grid, long M) {
    const double s0 = ph.stdv[0], s1 = ph.stdv[1], s2 = ph.stdv[2];
    const double m0 = (double)ph.mean[0] + (double)ph.depth, m1 = ph.mean[1], m2 = ph.mean[2], g = ph.g;
    float4 c[K];
    bool in[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int i = threadIdx.x + CPT * k;
        in[k] = i < S.n;
        c[k] = in[k] ? x[i] : make_float4(0.f, 0.f, 0.f, 1.f);
    }
    double v[2] = {0, 0};
#pragma unroll
    for (int k = 0; k < K; k++)
        if (in[k]) {
            v[0] += (double)c[k].x;
            v[1] += c[k].w != 0.f ? 0.0 : 1.0;
        }
    block_sum<2, CPT>(v, sh);
    const double nb = v[1];
    if (nb > 0) {
        const double delta = (mass * (double)S.n - v[0]) / nb;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (c[k].w == 0.f) c[k].x = (float)((double)c[k].x + delta);
        double e[5] = {0, 0, 0, 0, 0};  // pk, kk, pb, kb, sum of blank h
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!in[k]) continue;
            double h, u, w;
            phys_huv(c[k], ph, h, u, w);
            const double pp = 0.5 * g * h * h, kq = 0.5 * h * (u * u + w * w);
            if (c[k].w != 0.f) {
                e[0] += pp;
                e[1] += kq;
            } else {
                e[2] += pp;
                e[3] += kq;
                e[4] += h;
            }
        }
        block_sum<5, CPT>(e, sh);
        const double bound = (double)ph.r_max * eref;
        if (e[0] + e[1] + e[2] + e[3] > bound) {
            const double target = bound * (1.0 - 1e-7);
            const double budget = target - e[0] - e[1] - e[2];
            double su = 0.0;
            if (budget >= 0) su = fmin(sqrt(budget / fmax(e[3], 1e-12)), 1.0);
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (c[k].w != 0.f) continue;
                const double u = (double)c[k].y * s1 + m1, w = (double)c[k].z * s2 + m2;
                c[k].y = (float)((u * su - m1) / s1);
                c[k].z = (float)((w * su - m2) / s2);
            }
            if (budget < 0) {
                const double hbar = e[4] / nb;
                double a[1] = {0};
#pragma unroll
                for (int k = 0; k < K; k++) {
                    if (c[k].w != 0.f) continue;
                    const double d = (double)c[k].x * s0 + m0 - hbar;
                    a[0] += d * d;
                }
                block_sum<1, CPT>(a, sh);
                const double avail = target - e[0] - e[1] - 0.5 * g * nb * hbar * hbar;
                double lam = 0.0;
                if (avail > 0 && a[0] > 0) lam = fmin(sqrt(avail / (0.5 * g * a[0])), 1.0);
#pragma unroll
                for (int k = 0; k < K; k++) {
                    if (c[k].w != 0.f) continue;
                    const double h = (double)c[k].x * s0 + m0;
                    c[k].x = (float)(((hbar + lam * (h - hbar)) - m0) / s0);
                }
            }
        }
    }
    if (viol) {  // k_cell_violation's sum, in its order, on the final values
        double v[1] = {0};
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!in[k]) continue;
            double h, u, w;
            phys_huv(c[k], ph, h, u, w);
            v[0] += 0.5 * (double)ph.g * h * h + 0.5 * h * (u * u + w * w);
        }
        block_sum<1, CPT>(v, sh);
        if (threadIdx.x == 0)
            atomicMax(viol + S.req, (unsigned long long)__double_as_longlong(fmax((v[0] - (double)ph.r_max * eref) / eref, 0.0)));
    }
    float r = 0.f;
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (!in[k]) continue;
        x[threadIdx.x + CPT * k] = c[k];
        if (grid) {
            const long m = (long)S.start + threadIdx.x + CPT * k;
            grid[m] = c[k].x;
            grid[M + m] = c[k].y;
            grid[2 * M + m] = c[k].z;
        }
        const float4 b = p[threadIdx.x + CPT * k];
        if (c[k].w != 0.f) continue;
        r = fmaxf(r, fmaxf(fabsf(c[k].x - b.x), fmaxf(fabsf(c[k].y - b.y), fabsf(c[k].z - b.z))));
    }
    r = wave_max(r);
    if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w = 0; w < CPT / 64; w++) t = fmaxf(t, shm[w]);
        atomicMax(resid + S.req, __float_as_uint(t));
    }
}

// Blank-cell projection of one slice: mass shift, then scale blank (u, v)
// down to r_max*Eref*(1-1e-7); if potential alone is over it, zero blank
// (u, v) and shrink blank h anomalies about their mean (mass unchanged).
// Then the slice's max |xn - xc| over blank cells. Slices of up to 4096
// cells keep their cells in registers (cell_project_regs). With grid, the
// projected cells are also written to grid [3][M] (the unpack).
__global__ __launch_bounds__(CPT) void k_cell_project(const Slice* __restrict__ sl, float4* __restrict__ xn,
                                                     const float4* __restrict__ xc, const double* __restrict__ mass_t,
                                                     const double* __restrict__ eref_t, unsigned* __restrict__ resid,
                                                     unsigned long long* __restrict__ viol, Phys ph,
                                                     float* __restrict__ grid, long M) {
    __shared__ double sh[5][CPT / 64];
    __shared__ float shm[CPT / 64];
    const Slice S = sl[blockIdx.x];
    float4* x = xn + S.start;
    if (S.n <= 4 * CPT) {
        cell_project_regs<4>(S, x, xc + S.start, mass_t[blockIdx.x], eref_t[blockIdx.x], resid, viol, ph, sh, shm,
                             grid, M);
        return;
    }
    const double s0 = ph.stdv[0], s1 = ph.stdv[1], s2 = ph.stdv[2];
    const double m0 = (double)ph.mean[0] + (double)ph.depth, m1 = ph.mean[1], m2 = ph.mean[2], g = ph.g;
    {
        double v[2] = {0, 0};
        for (int i = threadIdx.x; i < S.n; i += CPT) {
            const float4 c = x[i];
            v[0] += (double)c.x;
            v[1] += c.w != 0.f ? 0.0 : 1.0;
        }
        block_sum<2, CPT>(v, sh);
        const double nb = v[1];
        if (nb > 0) {
            const double delta = (mass_t[blockIdx.x] * (double)S.n - v[0]) / nb;
            for (int i = threadIdx.x; i < S.n; i += CPT)
                if (x[i].w == 0.f) x[i].x = (float)((double)x[i].x + delta);
            __syncthreads();
            double e[5] = {0, 0, 0, 0, 0};  // pk, kk, pb, kb, sum of blank h
            for (int i = threadIdx.x; i < S.n; i += CPT) {
                const float4 c = x[i];
                double h, u, w;
                phys_huv(c, ph, h, u, w);
                const double pp = 0.5 * g * h * h, kq = 0.5 * h * (u * u + w * w);
                if (c.w != 0.f) {
                    e[0] += pp;
                    e[1] += kq;
                } else {
                    e[2] += pp;
                    e[3] += kq;
                    e[4] += h;
                }
            }
            block_sum<5, CPT>(e, sh);
            const double bound = (double)ph.r_max * eref_t[blockIdx.x];
            if (e[0] + e[1] + e[2] + e[3] > bound) {
                const double target = bound * (1.0 - 1e-7);
                const double budget = target - e[0] - e[1] - e[2];
                double su = 0.0;
                if (budget >= 0) su = fmin(sqrt(budget / fmax(e[3], 1e-12)), 1.0);
                for (int i = threadIdx.x; i < S.n; i += CPT) {
                    float4 c = x[i];
                    if (c.w != 0.f) continue;
                    const double u = (double)c.y * s1 + m1, w = (double)c.z * s2 + m2;
                    c.y = (float)((u * su - m1) / s1);
                    c.z = (float)((w * su - m2) / s2);
                    x[i] = c;
                }
                if (budget < 0) {
                    const double hbar = e[4] / nb;
                    double a[1] = {0};
                    for (int i = threadIdx.x; i < S.n; i += CPT) {
                        const float4 c = x[i];
                        if (c.w != 0.f) continue;
                        const double d = (double)c.x * s0 + m0 - hbar;
                        a[0] += d * d;
                    }
                    block_sum<1, CPT>(a, sh);
                    const double avail = target - e[0] - e[1] - 0.5 * g * nb * hbar * hbar;
                    double lam = 0.0;
                    if (avail > 0 && a[0] > 0) lam = fmin(sqrt(avail / (0.5 * g * a[0])), 1.0);
                    for (int i = threadIdx.x; i < S.n; i += CPT) {
                        float4 c = x[i];
                        if (c.w != 0.f) continue;
                        const double h = (double)c.x * s0 + m0;
                        c.x = (float)(((hbar + lam * (h - hbar)) - m0) / s0);
                        x[i] = c;
                    }
                }
                __syncthreads();
            }
        }
    }
    if (viol) {  // k_cell_violation's sum on the final values (each thread reads what it wrote)
        double v[1] = {0};
        for (int i = threadIdx.x; i < S.n; i += CPT) {
            double h, u, w;
            phys_huv(x[i], ph, h, u, w);
            v[0] += 0.5 * (double)ph.g * h * h + 0.5 * h * (u * u + w * w);
        }
        block_sum<1, CPT>(v, sh);
        if (threadIdx.x == 0) {
            const double e = eref_t[blockIdx.x];
            atomicMax(viol + S.req, (unsigned long long)__double_as_longlong(fmax((v[0] - (double)ph.r_max * e) / e, 0.0)));
        }
    }
    // Residual over blank cells (all three channels), max is order-free.
    float r = 0.f;
    const float4* p = xc + S.start;
    for (int i = threadIdx.x; i < S.n; i += CPT) {
        const float4 a = x[i], b = p[i];
        if (grid) {
            const long m = (long)S.start + i;
            grid[m] = a.x;
            grid[M + m] = a.y;
            grid[2 * M + m] = a.z;
        }
        if (a.w != 0.f) continue;
        r = fmaxf(r, fmaxf(fabsf(a.x - b.x), fmaxf(fabsf(a.y - b.y), fabsf(a.z - b.z))));
    }
    r = wave_max(r);
    if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w = 0; w < CPT / 64; w++) t = fmaxf(t, shm[w]);
        atomicMax(resid + S.req, __float_as_uint(t));
    }
}

void cell_project(hipStream_t s, int n_slices, const Slice* sl, float* xn, const float* xc, const double* mass_t,
                  const double* eref_t, unsigned* resid_bits, float g, const float* mean, const float* stdv,
                  float depth, float r_max, unsigned long long* viol_bits, float* grid, int M) {
    if (n_slices <= 0) return;
    hipLaunchKernelGGL(k_cell_project, dim3(n_slices), dim3(CPT), 0, s, sl, (float4*)xn, (const float4*)xc, mass_t,
                       eref_t, resid_bits, viol_bits, phys_of(g, mean, stdv, depth, r_max), grid, (long)M);
    LAUNCHCHK("cell_project");
}

__global__ __launch_bounds__(CPT) void k_cell_violation(const Slice* __restrict__ sl, const float4* __restrict__ x4,
                                                       const double* __restrict__ eref_t,
                                                       unsigned long long* __restrict__ viol, Phys ph,
                                                       const int* __restrict__ eref_idx) {
    __shared__ double sh[1][CPT / 64];
    const Slice S = sl[blockIdx.x];
    double v[1] = {0};
    for (int i = threadIdx.x; i < S.n; i += CPT) {
        double h, u, w;
        phys_huv(x4[S.start + i], ph, h, u, w);
        v[0] += 0.5 * (double)ph.g * h * h + 0.5 * h * (u * u + w * w);
    }
    block_sum<1, CPT>(v, sh);
    if (threadIdx.x == 0) {
        const double e = eref_t[eref_idx ? eref_idx[blockIdx.x] : (int)blockIdx.x];
        const double r = fmax((v[0] - (double)ph.r_max * e) / e, 0.0);
        atomicMax(viol + S.req, (unsigned long long)__double_as_longlong(r));
    }
}

void cell_violation(hipStream_t s, int n_slices, const Slice* sl, const float* x4, const double* eref_t,
                    unsigned long long* viol_bits, float g, const float* mean, const float* stdv, float depth,
                    float r_max, const int* eref_idx) {
    if (n_slices <= 0) return;
    hipLaunchKernelGGL(k_cell_violation, dim3(n_slices), dim3(CPT), 0, s, sl, (const float4*)x4, eref_t, viol_bits,
                       phys_of(g, mean, stdv, depth, r_max), eref_idx);
    LAUNCHCHK("cell_violation");
}

}  // namespace j2k
