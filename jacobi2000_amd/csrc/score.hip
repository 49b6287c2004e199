// Field scores on the device: crates/engine/src/metrics.rs (correlation,
// peak, mean, std_dev, bias, radial_profile, decode_h) as eval-metrics
// applies them to a model rollout against the solver
// (crates/engine-cli/src/eval.rs:158-172), and the rollout CLI's
// per-step correlation (crates/engine-cli/src/main.rs:123-125).
//
// Arithmetic: every per-cell value (encode, decode_h, |x|, x - mean, the
// squares, the radius and its bin) is the reference's f32 expression with
// contraction off (the pragma below and -ffp-contract=off on this file).
// Sums run in f64, in a fixed order (a per-thread stride, then a fixed
// butterfly and tree), so a score does not change between runs. The
// reference sums correlation terms in f64 left to right, so the two differ
// by f64 rounding only; its mean and std_dev sum in f32, which the f64 sums
// here carry more exactly. Radial bin sums are f64 atomics (their order, at
// the 1e-16 level, is the one thing that varies). The host finishes the
// O(bins x lags) wavefront lag (metrics.rs:63-107) as the reference does.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "abi.h"
#include "model.h"
#include "score.h"

#pragma clang fp contract(off)

namespace j2k {
namespace {

constexpr int SNT = 256;  // threads per scoring block

// Sum over the block in a fixed order: a wave butterfly, then wave totals
// added in wave order by thread 0. Every thread gets the total.
__device__ double block_sum(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < SNT / 64; i++) t += red[i];
    return t;
}

__device__ float block_max(float v, float* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    float t = 0.0f;
#pragma unroll
    for (int i = 0; i < SNT / 64; i++) t = fmaxf(t, red[i]);
    return t;
}

// metrics::correlation of (a*sa + ta) against (b*sb + tb) over n values
// (metrics.rs:9-23): means, then the three sums of deviations; 0 when
// either field is constant.
__global__ __launch_bounds__(SNT) void k_corr_affine(const float* a, float sa, float ta, const float* b, float sb,
                                                     float tb, int n, float* out) {
    __shared__ double red[SNT / 64];
    double s1 = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < n; i += SNT) {
        s1 += (double)(a[i] * sa + ta);
        s2 += (double)(b[i] * sb + tb);
    }
    const double nn = (double)n;
    const double ma = block_sum(s1, red) / nn;
    const double mb = block_sum(s2, red) / nn;
    double sab = 0.0, saa = 0.0, sbb = 0.0;
    for (int i = threadIdx.x; i < n; i += SNT) {
        const double dx = (double)(a[i] * sa + ta) - ma;
        const double dy = (double)(b[i] * sb + tb) - mb;
        sab += dx * dy;
        saa += dx * dx;
        sbb += dy * dy;
    }
    sab = block_sum(sab, red);
    saa = block_sum(saa, red);
    sbb = block_sum(sbb, red);
    if (threadIdx.x == 0) *out = (saa <= 0.0 || sbb <= 0.0) ? 0.0f : (float)(sab / sqrt(saa * sbb));
}

struct ScoreNorm {
    float mean[3], std[3], depth;
};

// Per-trajectory sums of the eval-metrics scores: slots of sums[b][].
enum : int {
    SC_SAB, SC_SAA, SC_SBB,        // correlation over (h, u, v), normalized fields
    SC_SAB_H, SC_SAA_H, SC_SBB_H,  // over h
    SC_MEAN_P, SC_MEAN_T,          // mean of decoded model h / solver h (f32 values)
    SC_VAR_T,                      // sum of (th - mean_t)^2
    SC_PEAK_P, SC_PEAK_T,          // peak |h|
    SC_RMAX,                       // last radial bin
    SC_COUNT
};

// One block per trajectory b: x [3n] normalized model state, (h, u, v) [n]
// the solver's physical state, (cx, cy) the drop centre. Writes the sums to
// sums[b][SC_COUNT] and the radial profiles of |decoded h| to
// prof[b][2][nbins] (model, solver; metrics.rs:41-61).
__global__ __launch_bounds__(SNT) void k_score(const float* xs, const float* hs, const float* us, const float* vs,
                                               ScoreNorm nm, int nx, int ny, const float* centers, int nbins,
                                               double* sums, double* prof) {
    extern __shared__ double bins[];  // [2][nbins] sums, then nbins counts
    __shared__ double red[SNT / 64];
    __shared__ float redf[SNT / 64];
    unsigned* cnt = reinterpret_cast<unsigned*>(bins + 2 * nbins);
    const int b = blockIdx.x;
    const int n = nx * ny;
    const float* x = xs + (size_t)b * 3 * n;
    const float* h = hs + (size_t)b * n;
    const float* u = us + (size_t)b * n;
    const float* v = vs + (size_t)b * n;
    const float cx = centers[2 * b], cy = centers[2 * b + 1];
    // metrics.rs:44: the last bin, cut at the nearest wall
    const float rm = fmaxf(floorf(fminf(fminf(fminf(cx, cy), (float)nx - cx), (float)ny - cy)), 1.0f);
    const int rmax = min((int)rm, nbins - 1);
    for (int i = threadIdx.x; i < 2 * nbins; i += SNT) bins[i] = 0.0;
    for (int i = threadIdx.x; i < nbins; i += SNT) cnt[i] = 0u;
    __syncthreads();

    // norm.rs:27-34 encode of the solver state; metrics.rs:111-113 decode_h
    auto target = [&](int c, int i) -> float {
        if (c == 0) return (h[i] - nm.depth - nm.mean[0]) / nm.std[0];
        if (c == 1) return (u[i] - nm.mean[1]) / nm.std[1];
        return (v[i] - nm.mean[2]) / nm.std[2];
    };
    double sx = 0.0, st = 0.0, sxh = 0.0, sth = 0.0, sp = 0.0, sq = 0.0;
    float pk_p = 0.0f, pk_t = 0.0f;
    for (int i = threadIdx.x; i < n; i += SNT) {
        const float x0 = x[i], t0 = target(0, i);
        const float x1 = x[n + i], t1 = target(1, i);
        const float x2 = x[2 * n + i], t2 = target(2, i);
        sx += (double)x0;
        sx += (double)x1;
        sx += (double)x2;
        st += (double)t0;
        st += (double)t1;
        st += (double)t2;
        sxh += (double)x0;
        sth += (double)t0;
        const float ph = x0 * nm.std[0] + nm.mean[0];
        const float th = t0 * nm.std[0] + nm.mean[0];
        sp += (double)ph;
        sq += (double)th;
        pk_p = fmaxf(pk_p, fabsf(ph));
        pk_t = fmaxf(pk_t, fabsf(th));
        const int yy = i / nx, xx = i - yy * nx;
        const float dx = (float)xx - cx, dy = (float)yy - cy;
        const float r = sqrtf(dx * dx + dy * dy);
        const int bin = (int)(r + 0.5f);
        if (bin <= rmax) {
            atomicAdd(&bins[bin], (double)fabsf(ph));
            atomicAdd(&bins[nbins + bin], (double)fabsf(th));
            atomicAdd(&cnt[bin], 1u);
        }
    }
    const double ma = block_sum(sx, red) / (3.0 * n);
    const double mb = block_sum(st, red) / (3.0 * n);
    const double mah = block_sum(sxh, red) / (double)n;
    const double mbh = block_sum(sth, red) / (double)n;
    // metrics.rs:29-31: mean = sum / len in f32
    const float mean_p = (float)block_sum(sp, red) / (float)n;
    const float mean_t = (float)block_sum(sq, red) / (float)n;
    pk_p = block_max(pk_p, redf);
    pk_t = block_max(pk_t, redf);

    double sab = 0.0, saa = 0.0, sbb = 0.0, sabh = 0.0, saah = 0.0, sbbh = 0.0, svt = 0.0;
    for (int i = threadIdx.x; i < n; i += SNT) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double dx = (double)x[c * n + i] - ma;
            const double dy = (double)target(c, i) - mb;
            sab += dx * dy;
            saa += dx * dx;
            sbb += dy * dy;
        }
        const double dxh = (double)x[i] - mah;
        const double dyh = (double)target(0, i) - mbh;
        sabh += dxh * dyh;
        saah += dxh * dxh;
        sbbh += dyh * dyh;
        // metrics.rs:33-36: (x - m)^2 in f32
        const float th = target(0, i) * nm.std[0] + nm.mean[0];
        const float d = th - mean_t;
        svt += (double)(d * d);
    }
    double out[SC_COUNT];
    out[SC_SAB] = block_sum(sab, red);
    out[SC_SAA] = block_sum(saa, red);
    out[SC_SBB] = block_sum(sbb, red);
    out[SC_SAB_H] = block_sum(sabh, red);
    out[SC_SAA_H] = block_sum(saah, red);
    out[SC_SBB_H] = block_sum(sbbh, red);
    out[SC_MEAN_P] = mean_p;
    out[SC_MEAN_T] = mean_t;
    out[SC_VAR_T] = block_sum(svt, red);
    out[SC_PEAK_P] = pk_p;
    out[SC_PEAK_T] = pk_t;
    out[SC_RMAX] = rmax;
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < SC_COUNT; j++) sums[(size_t)b * SC_COUNT + j] = out[j];
    __syncthreads();  // the bins are complete
    for (int j = threadIdx.x; j < nbins; j += SNT) {
        const unsigned c = cnt[j];
        prof[((size_t)b * 2) * nbins + j] = c == 0 ? 0.0 : bins[j] / (double)c;
        prof[((size_t)b * 2 + 1) * nbins + j] = c == 0 ? 0.0 : bins[nbins + j] / (double)c;
    }
}

// ---- host: the scalar finish (metrics.rs) --------------------------------------

// metrics.rs:9-23 on host values (the wavefront lag correlates short
// profile slices).
float correlation_host(const std::vector<float>& a, const std::vector<float>& b) {
    const double n = (double)a.size();
    double ma = 0.0, mb = 0.0;
    for (float x : a) ma += (double)x;
    for (float y : b) mb += (double)y;
    ma /= n;
    mb /= n;
    double sab = 0.0, saa = 0.0, sbb = 0.0;
    for (size_t i = 0; i < a.size(); i++) {
        const double dx = (double)a[i] - ma, dy = (double)b[i] - mb;
        sab += dx * dy;
        saa += dx * dx;
        sbb += dy * dy;
    }
    if (saa <= 0.0 || sbb <= 0.0) return 0.0f;
    return (float)(sab / std::sqrt(saa * sbb));
}

// metrics.rs:63-76
float corr_at_lag(const double* model, const double* solver, int n, int lag) {
    std::vector<float> a, b;
    for (int i = 0; i < n; i++) {
        const int si = i - lag;
        if (si >= 0 && si < n) {
            a.push_back((float)model[i]);
            b.push_back((float)solver[si]);
        }
    }
    if (a.size() < 3) return -std::numeric_limits<float>::infinity();
    return correlation_host(a, b);
}

}  // namespace

// metrics.rs:78-107: the lag maximising the profile cross-correlation,
// refined by a parabola through its neighbours.
float xcorr_lag_host(const double* model, const double* solver, int n, int max_lag) {
    int best_lag = 0;
    float best_c = -std::numeric_limits<float>::infinity();
    for (int lag = -max_lag; lag <= max_lag; lag++) {
        const float c = corr_at_lag(model, solver, n, lag);
        if (c > best_c) {
            best_c = c;
            best_lag = lag;
        }
    }
    if (best_lag > -max_lag && best_lag < max_lag) {
        const float l = corr_at_lag(model, solver, n, best_lag - 1), c = best_c,
                    r = corr_at_lag(model, solver, n, best_lag + 1);
        if (std::isfinite(l) && std::isfinite(r)) {
            const float den = l - 2.0f * c + r;
            if (den < 0.0f) return (float)best_lag + 0.5f * (l - r) / den;
        }
    }
    return (float)best_lag;
}

void corr_affine(hipStream_t s, const float* a, float sa, float ta, const float* b, float sb, float tb, int n,
                 float* out) {
    hipLaunchKernelGGL(k_corr_affine, dim3(1), dim3(SNT), 0, s, a, sa, ta, b, sb, tb, n, out);
    LAUNCHCHK("k_corr_affine");
}

// The radial profile's bin count for a batch: every trajectory's last bin
// (metrics.rs:44) plus one.
int score_bins(const float* centers, int batch, int nx, int ny) {
    int nb = 2;
    for (int b = 0; b < batch; b++) {
        const float cx = centers[2 * b], cy = centers[2 * b + 1];
        const float rm = std::max(std::floor(std::min(std::min(std::min(cx, cy), (float)nx - cx), (float)ny - cy)), 1.0f);
        nb = std::max(nb, (int)rm + 1);
    }
    return nb;
}

void score_batch(Model& m, hipStream_t s, int batch, const float* d_x, const float* d_h, const float* d_u,
                 const float* d_v, const float* centers, int max_lag, j2k_eval_score* out) {
    const j2k_config& c = m.cfg;
    const int nx = c.nx, ny = c.ny;
    for (int b = 0; b < batch; b++)
        if (!(centers[2 * b] == centers[2 * b]) || !(centers[2 * b + 1] == centers[2 * b + 1]))
            fail(J2K_ERR_INVALID, "score: a drop centre is NaN");
    const int nbins = score_bins(centers, batch, nx, ny);
    if ((size_t)nbins * 20 > 64 * 1024) fail(J2K_ERR_LIMIT, "score: too many radial bins for one block's LDS");
    // one block: sums [batch][SC_COUNT], profiles [batch][2][nbins], centres [batch][2] (floats)
    const size_t ns = (size_t)batch * SC_COUNT, np = (size_t)batch * 2 * nbins;
    double* d_sums = m.get<double>(S_SUMS, ns + np + (size_t)batch);
    double* d_prof = d_sums + ns;
    float* d_c = reinterpret_cast<float*>(d_prof + np);
    HIPCHK(hipMemcpyAsync(d_c, centers, (size_t)2 * batch * sizeof(float), hipMemcpyHostToDevice, s));
    ScoreNorm nm;
    for (int i = 0; i < 3; i++) nm.mean[i] = c.norm_mean[i], nm.std[i] = c.norm_std[i];
    nm.depth = c.norm_depth;
    hipLaunchKernelGGL(k_score, dim3(batch), dim3(SNT), (size_t)nbins * 20, s, d_x, d_h, d_u, d_v, nm, nx, ny, d_c,
                       nbins, d_sums, d_prof);
    LAUNCHCHK("k_score");
    std::vector<double> sums(ns), prof(np);
    HIPCHK(hipMemcpyAsync(sums.data(), d_sums, ns * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(prof.data(), d_prof, np * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const int n = nx * ny;
    for (int b = 0; b < batch; b++) {
        const double* q = &sums[(size_t)b * SC_COUNT];
        auto corr = [](double sab, double saa, double sbb) {
            return (saa <= 0.0 || sbb <= 0.0) ? 0.0f : (float)(sab / std::sqrt(saa * sbb));
        };
        j2k_eval_score& o = out[b];
        o.corr = corr(q[SC_SAB], q[SC_SAA], q[SC_SBB]);
        o.corr_h = corr(q[SC_SAB_H], q[SC_SAA_H], q[SC_SBB_H]);
        o.amplitude_ratio = (float)q[SC_PEAK_P] / (float)q[SC_PEAK_T];
        const int bins = (int)q[SC_RMAX] + 1;
        o.lag = xcorr_lag_host(&prof[(size_t)b * 2 * nbins], &prof[((size_t)b * 2 + 1) * nbins], bins, max_lag);
        // metrics.rs:33-40
        const float std_t = std::sqrt((float)q[SC_VAR_T] / (float)n);
        o.bias = ((float)q[SC_MEAN_P] - (float)q[SC_MEAN_T]) / std::max(std_t, 1e-12f);
    }
}

}  // namespace j2k

using namespace j2k;

extern "C" {

int j2k_correlation_dev(const float* d_a, float sa, float ta, const float* d_b, float sb, float tb, int n,
                        float* d_out, void* stream) {
    return guard([&] {
        if (!d_a || !d_b || !d_out) fail(J2K_ERR_INVALID, "j2k_correlation_dev: NULL pointer");
        if (n < 1) fail(J2K_ERR_INVALID, "j2k_correlation_dev: n must be >= 1");
        corr_affine(static_cast<hipStream_t>(stream), d_a, sa, ta, d_b, sb, tb, n, d_out);
    });
}

}  // extern "C"
