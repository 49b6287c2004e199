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

// The blank-cell projection of one slice (mass, then energy) is planned
// from eight f64 moments of the slice taken in one pass over the cells it
// holds before the projection: over all cells the sum of normalized h;
// over known cells the count and the potential and kinetic energy sums;
// over blank cells the count and the sums of physical h, h^2, q = u^2 + v^2
// and h q. The mass shift delta (normalized) moves blank h by ds = delta s0,
// so after it sum h = Sh + nb ds, sum h^2 = Sh2 + 2 ds Sh + nb ds^2 and
// sum h q = Shq + ds Sq give the blank energies without a second pass; the
// anomaly sum of the shrink case is sum h^2 - nb hbar^2. One block
// reduction plans the projection instead of one per pass.
enum { MO_X, MO_NB, MO_PK, MO_KK, MO_H, MO_H2, MO_Q, MO_HQ, MO_N };

struct ProjPlan {
    double delta = 0, su = 1, hbar = 0, lam = 1;
    bool shift = false, scale = false, shrink = false;
};

__device__ __forceinline__ void add_moments(double (&mo)[MO_N], const float4& c, const Phys& ph) {
    double h, u, w;
    phys_huv(c, ph, h, u, w);
    const double q = u * u + w * w;
    // weights instead of branches: a branch here becomes a select of the
    // array index, which puts the moments in scratch memory
    const double kn = c.w != 0.f ? 1.0 : 0.0, bl = 1.0 - kn;
    mo[MO_X] += (double)c.x;
    mo[MO_PK] += kn * (0.5 * (double)ph.g * h * h);
    mo[MO_KK] += kn * (0.5 * h * q);
    mo[MO_NB] += bl;
    mo[MO_H] += bl * h;
    mo[MO_H2] += bl * (h * h);
    mo[MO_Q] += bl * q;
    mo[MO_HQ] += bl * (h * q);
}

__device__ __forceinline__ ProjPlan plan_projection(const double (&mo)[MO_N], int n, double mass, double eref,
                                                    const Phys& ph) {
    ProjPlan pl;
    const double nb = mo[MO_NB];
    if (!(nb > 0)) return pl;
    const double g = ph.g;
    pl.shift = true;
    pl.delta = (mass * (double)n - mo[MO_X]) / nb;
    const double ds = pl.delta * (double)ph.stdv[0];
    const double sh = mo[MO_H] + nb * ds, sh2 = mo[MO_H2] + 2.0 * ds * mo[MO_H] + nb * ds * ds;
    const double pb = 0.5 * g * sh2, kb = 0.5 * (mo[MO_HQ] + ds * mo[MO_Q]);
    const double bound = (double)ph.r_max * eref;
    if (mo[MO_PK] + mo[MO_KK] + pb + kb > bound) {
        const double target = bound * (1.0 - 1e-7);
        const double budget = target - mo[MO_PK] - mo[MO_KK] - pb;
        pl.scale = true;
        pl.su = budget >= 0 ? fmin(sqrt(budget / fmax(kb, 1e-12)), 1.0) : 0.0;
        if (budget < 0) {
            pl.shrink = true;
            pl.hbar = sh / nb;
            const double a = fmax(sh2 - nb * pl.hbar * pl.hbar, 0.0);
            const double avail = target - mo[MO_PK] - mo[MO_KK] - 0.5 * g * nb * pl.hbar * pl.hbar;
            pl.lam = (avail > 0 && a > 0) ? fmin(sqrt(avail / (0.5 * g * a)), 1.0) : 0.0;
        }
    }
    return pl;
}

// The plan applied to one cell (blank cells only): shift h, then scale
// (u, v), then shrink the h anomaly about hbar; float rounding after each.
__device__ __forceinline__ void apply_projection(float4& c, const ProjPlan& pl, const Phys& ph) {
    if (c.w != 0.f) return;
    const double s0 = ph.stdv[0], s1 = ph.stdv[1], s2 = ph.stdv[2];
    const double m0 = (double)ph.mean[0] + (double)ph.depth, m1 = ph.mean[1], m2 = ph.mean[2];
    if (pl.shift) c.x = (float)((double)c.x + pl.delta);
    if (pl.scale) {
        const double u = (double)c.y * s1 + m1, w = (double)c.z * s2 + m2;
        c.y = (float)((u * pl.su - m1) / s1);
        c.z = (float)((w * pl.su - m2) / s2);
    }
    if (pl.shrink) {
        const double h = (double)c.x * s0 + m0;
        c.x = (float)(((pl.hbar + pl.lam * (h - pl.hbar)) - m0) / s0);
    }
}

// Block totals of the moments, every thread gets them: wave sums, then one
// thread per moment adds the waves' partials in wave order.
__device__ __forceinline__ void moments_reduce(double (&mo)[MO_N], double (*sh)[CPT / 64], double* tot) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < MO_N; j++) {
        const double t = wave_sum_d(mo[j]);
        if (lane == 0) sh[j][wave] = t;
    }
    __syncthreads();
    if ((int)threadIdx.x < MO_N) {
        double t = 0;
        for (int w = 0; w < CPT / 64; w++) t += sh[threadIdx.x][w];
        tot[threadIdx.x] = t;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < MO_N; j++) mo[j] = tot[j];
}

// Block totals of the energy of the projected cells (the violation sum) and
// of the residual max, in one round: thread 0 gets both.
__device__ __forceinline__ void viol_resid_reduce(double& v, float& r, double* shv, float* shr) {
    v = wave_sum_d(v);
    r = wave_max(r);
    if ((threadIdx.x & 63) == 0) {
        shv[threadIdx.x >> 6] = v;
        shr[threadIdx.x >> 6] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tv = 0;
        float tr = 0.f;
        for (int w = 0; w < CPT / 64; w++) {
            tv += shv[w];
            tr = fmaxf(tr, shr[w]);
        }
        v = tv;
        r = tr;
    }
}

__device__ __forceinline__ void finish_slice(double v, float r, const Slice& S, double eref, const Phys& ph,
                                             unsigned* __restrict__ resid, unsigned long long* __restrict__ viol) {
    if (threadIdx.x != 0) return;
    if (viol)
        atomicMax(viol + S.req,
                  (unsigned long long)__double_as_longlong(fmax((v - (double)ph.r_max * eref) / eref, 0.0)));
    atomicMax(resid + S.req, __float_as_uint(r));
}

// Blank-cell projection of one slice, planned from its moments (above):
// mass shift, then scale blank (u, v) down to r_max*Eref*(1-1e-7); if
// potential alone is over it, zero blank (u, v) and shrink blank h
// anomalies about their mean (mass unchanged). Then the slice's energy on
// the projected cells (the violation) and max |xn - xc| over blank cells.
// Slices of up to 4096 cells keep their cells in registers from the one
// load to the one store; larger ones take two passes over memory. With
// grid, the projected cells are also written to grid [3][M] (the unpack).
__global__ __launch_bounds__(CPT) void k_cell_project(const Slice* __restrict__ sl, float4* __restrict__ xn,
                                                     const float4* __restrict__ xc, const double* __restrict__ mass_t,
                                                     const double* __restrict__ eref_t, unsigned* __restrict__ resid,
                                                     unsigned long long* __restrict__ viol, Phys ph,
                                                     float* __restrict__ grid, long M) {
    __shared__ double sh[MO_N][CPT / 64];
    __shared__ double tot[MO_N];
    __shared__ float shm[CPT / 64];
    const Slice S = sl[blockIdx.x];
    float4* x = xn + S.start;
    const float4* p = xc + S.start;
    const double mass = mass_t[blockIdx.x], eref = eref_t[blockIdx.x];
    auto store = [&](int i, const float4& c) {
        x[i] = c;
        if (grid) {
            const long m = (long)S.start + i;
            grid[m] = c.x;
            grid[M + m] = c.y;
            grid[2 * M + m] = c.z;
        }
    };
    auto energy = [&](const float4& c) {
        double h, u, w;
        phys_huv(c, ph, h, u, w);
        return 0.5 * (double)ph.g * h * h + 0.5 * h * (u * u + w * w);
    };
    double mo[MO_N] = {0, 0, 0, 0, 0, 0, 0, 0};
    constexpr int K = 4;
    if (S.n <= K * CPT) {
        float4 c[K], b[K];  // the cells, and the previous iterate's (read now, used for the residual)
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = threadIdx.x + CPT * k;
            if (i < S.n) {
                c[k] = x[i];
                b[k] = p[i];
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++)
            if (threadIdx.x + CPT * k < S.n) add_moments(mo, c[k], ph);
        moments_reduce(mo, sh, tot);
        const ProjPlan pl = plan_projection(mo, S.n, mass, eref, ph);
        double v = 0;
        float r = 0.f;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = threadIdx.x + CPT * k;
            if (i >= S.n) continue;
            apply_projection(c[k], pl, ph);
            store(i, c[k]);
            if (viol) v += energy(c[k]);
            if (c[k].w != 0.f) continue;
            r = fmaxf(r, fmaxf(fabsf(c[k].x - b[k].x), fmaxf(fabsf(c[k].y - b[k].y), fabsf(c[k].z - b[k].z))));
        }
        viol_resid_reduce(v, r, sh[0], shm);
        finish_slice(v, r, S, eref, ph, resid, viol);
        return;
    }
    for (int i = threadIdx.x; i < S.n; i += CPT) add_moments(mo, x[i], ph);
    moments_reduce(mo, sh, tot);
    const ProjPlan pl = plan_projection(mo, S.n, mass, eref, ph);
    double v = 0;
    float r = 0.f;
    for (int i = threadIdx.x; i < S.n; i += CPT) {
        float4 c = x[i];
        apply_projection(c, pl, ph);
        store(i, c);
        if (viol) v += energy(c);
        if (c.w != 0.f) continue;
        const float4 b = p[i];
        r = fmaxf(r, fmaxf(fabsf(c.x - b.x), fmaxf(fabsf(c.y - b.y), fabsf(c.z - b.z))));
    }
    viol_resid_reduce(v, r, sh[0], shm);
    finish_slice(v, r, S, eref, ph, resid, viol);
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
