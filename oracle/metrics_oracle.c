/* CPU oracle for the field scores. TEST INFRASTRUCTURE ONLY: the checker for
 * libj2k's device scoring, never linked by it.
 *
 * Restates crates/engine/src/metrics.rs loop for loop (sums in the types and
 * order the reference uses: f64 for correlation and the radial profile, f32
 * for mean and std_dev) and the normalizer's encode (norm.rs:27-34). Pinned
 * in tests/test_oracle_metrics.py by the reference's metric tests
 * (metrics.rs:115-137) and by closed forms: the profile of radial fields,
 * integer lags of shifted profiles, the parabola refinement, bias of a
 * shifted field. */
#include <math.h>
#include <stdlib.h>

#include "j2k_oracle.h"

/* metrics.rs:25-27 */
float orc_peak(const float* f, int n) {
    float a = 0.0f;
    for (int i = 0; i < n; i++) a = fmaxf(a, fabsf(f[i]));
    return a;
}

/* metrics.rs:29-31 */
float orc_mean(const float* f, int n) {
    float s = 0.0f;
    for (int i = 0; i < n; i++) s += f[i];
    return s / (float)n;
}

/* metrics.rs:33-36 */
float orc_std_dev(const float* f, int n) {
    const float m = orc_mean(f, n);
    float s = 0.0f;
    for (int i = 0; i < n; i++) {
        const float d = f[i] - m;
        s += d * d;
    }
    return sqrtf(s / (float)n);
}

/* metrics.rs:38-41 */
float orc_bias(const float* model, const float* solver, int n) {
    return (orc_mean(model, n) - orc_mean(solver, n)) / fmaxf(orc_std_dev(solver, n), 1e-12f);
}

/* metrics.rs:43-61: out[0..rmax] (returns rmax + 1 bins). */
int orc_radial_profile(const float* f, int nx, int ny, float cx, float cy, double* out, int cap) {
    float rm = fminf(fminf(fminf(cx, cy), (float)nx - cx), (float)ny - cy);
    rm = fmaxf(floorf(rm), 1.0f);
    const int rmax = (int)rm;
    if (rmax + 1 > cap) return -1;
    double* sum = (double*)calloc((size_t)rmax + 1, sizeof(double));
    unsigned* count = (unsigned*)calloc((size_t)rmax + 1, sizeof(unsigned));
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const float dx = (float)x - cx, dy = (float)y - cy;
            const float r = sqrtf(dx * dx + dy * dy);
            const int bin = (int)(r + 0.5f);
            if (bin <= rmax) {
                sum[bin] += (double)fabsf(f[y * nx + x]);
                count[bin] += 1;
            }
        }
    for (int i = 0; i <= rmax; i++) out[i] = count[i] == 0 ? 0.0 : sum[i] / (double)count[i];
    free(sum);
    free(count);
    return rmax + 1;
}

/* metrics.rs:63-76 */
static float corr_at_lag(const double* model, const double* solver, int n, int lag) {
    float* a = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    float* b = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    int k = 0;
    for (int i = 0; i < n; i++) {
        const int si = i - lag;
        if (si >= 0 && si < n) {
            a[k] = (float)model[i];
            b[k] = (float)solver[si];
            k++;
        }
    }
    float c = -INFINITY;
    if (k >= 3) c = orc_correlation(a, b, k);
    free(a);
    free(b);
    return c;
}

/* metrics.rs:78-107 */
float orc_xcorr_lag(const double* model, const double* solver, int n, int max_lag) {
    int best_lag = 0;
    float best_c = -INFINITY;
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
        if (isfinite(l) && isfinite(r)) {
            const float den = l - 2.0f * c + r;
            if (den < 0.0f) return (float)best_lag + 0.5f * (l - r) / den;
        }
    }
    return (float)best_lag;
}

/* norm.rs:27-34: (h, u, v) [n] -> out [3][n] */
void orc_encode(const float* h, const float* u, const float* v, int n, const float* mean, const float* stdv,
                float depth, float* out) {
    for (int i = 0; i < n; i++) {
        out[i] = (h[i] - depth - mean[0]) / stdv[0];
        out[n + i] = (u[i] - mean[1]) / stdv[1];
        out[2 * n + i] = (v[i] - mean[2]) / stdv[2];
    }
}

/* metrics.rs:109-113 */
void orc_decode_h(const float* x, int n, const float* mean, const float* stdv, float* out) {
    for (int i = 0; i < n; i++) out[i] = x[i] * stdv[0] + mean[0];
}
