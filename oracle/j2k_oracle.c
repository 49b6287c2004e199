/* CPU oracle for the jacobi2000 engine forward pass. TEST INFRASTRUCTURE
 * ONLY: the checker for the HIP library, never linked by it. See
 * j2k_oracle.h for what each function restates and how it is pinned. */
#include "j2k_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* crates/engine/src/mask.rs:18-30. Rust's `%` and `/` on isize truncate
 * toward zero, as C's do. */
int orc_mask_allowed(int mask_kind, int radius, int dilation, int tokens_x, int q, int k) {
    if (mask_kind == 0) return 1;
    int qx = q % tokens_x, qy = q / tokens_x, kx = k % tokens_x, ky = k / tokens_x;
    int dx = kx - qx, dy = ky - qy, d = dilation;
    return dx % d == 0 && dy % d == 0 && abs(dx / d) <= radius && abs(dy / d) <= radius;
}

float orc_gelu_f32(float x) {
    const float inv_sqrt2 = 1.0f / sqrtf(2.0f);
    return 0.5f * x * (1.0f + erff(x * inv_sqrt2));
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* Form C blank fill. A blank cell takes the value of the same (y, x) cell in
 * the nearest fully known slice (ties go to the earlier slice). With no fully
 * known slice, it takes the mean of the known cells of its own slice in that
 * channel, else the mean of all known cells, else 0. Known cells keep theirs. */
void orc_fill_blank(const float* grid, const uint8_t* known, int T, int ny, int nx, float* out) {
    int n = ny * nx, M = T * n;
    memcpy(out, grid, sizeof(float) * 3 * (size_t)M);
    int* full = (int*)malloc(sizeof(int) * T);
    int nfull = 0;
    for (int s = 0; s < T; s++) {
        int all = 1;
        for (int i = 0; i < n && all; i++) all = known[(size_t)s * n + i] != 0;
        if (all) full[nfull++] = s;
    }
    double gsum[3] = {0, 0, 0};
    long gcnt = 0;
    for (int m = 0; m < M; m++)
        if (known[m]) {
            for (int c = 0; c < 3; c++) gsum[c] += grid[(size_t)c * M + m];
            gcnt++;
        }
    for (int s = 0; s < T; s++) {
        int src = -1, best = 1 << 30;
        for (int j = 0; j < nfull; j++) {
            int dd = abs(full[j] - s);
            if (dd < best) { best = dd; src = full[j]; }
        }
        float fillv[3] = {0, 0, 0};
        if (src < 0) {
            double ssum[3] = {0, 0, 0};
            long scnt = 0;
            for (int i = 0; i < n; i++)
                if (known[(size_t)s * n + i]) {
                    for (int c = 0; c < 3; c++) ssum[c] += grid[(size_t)c * M + (size_t)s * n + i];
                    scnt++;
                }
            for (int c = 0; c < 3; c++)
                fillv[c] = scnt ? (float)(ssum[c] / scnt) : (gcnt ? (float)(gsum[c] / gcnt) : 0.0f);
        }
        for (int i = 0; i < n; i++) {
            size_t m = (size_t)s * n + i;
            if (known[m]) continue;
            for (int c = 0; c < 3; c++)
                out[(size_t)c * M + m] = src >= 0 ? grid[(size_t)c * M + (size_t)src * n + i] : fillv[c];
        }
    }
    free(full);
}

static void slice_energy_f64(const float* x, const orc_phys* d, int M, int s, int n, double* mass, double* e) {
    double g = d->g, sh = 0, en = 0;
    for (int i = 0; i < n; i++) {
        size_t m = (size_t)s * n + i;
        double h = (double)x[m] * d->std[0] + ((double)d->mean[0] + (double)d->depth);
        double u = (double)x[(size_t)M + m] * d->std[1] + d->mean[1];
        double v = (double)x[2 * (size_t)M + m] * d->std[2] + d->mean[2];
        sh += (double)x[m];
        en += 0.5 * g * h * h + 0.5 * h * (u * u + v * v);
    }
    *mass = sh / (double)n;
    *e = en;
}

void orc_slice_targets_f64(const float* x0, const orc_phys* p, int T, int n, double* mass_t, double* eref_t) {
    for (int s = 0; s < T; s++) slice_energy_f64(x0, p, T * n, s, n, &mass_t[s], &eref_t[s]);
}

double orc_energy_violation_c(const float* x, const orc_phys* p, int T, int n, const double* eref_t) {
    double worst = 0;
    for (int s = 0; s < T; s++) {
        double m, e;
        slice_energy_f64(x, p, T * n, s, n, &m, &e);
        double v = (e - (double)p->r_max * eref_t[s]) / eref_t[s];
        if (v > worst) worst = v;
    }
    return worst;
}

/* bf16 activations: a mask of the values the trunk rounds to bf16 (nearest
 * even, through f32), the points where the served bf16 form stores bf16:
 * ORC_BF16_LN1 (LN1 output), ORC_BF16_QKV (q, k, v), ORC_BF16_ATTN (the
 * attention output), ORC_BF16_LN2 (LN2 output), ORC_BF16_GELU (GELU output). */
static int g_act_bf16 = 0;
void orc_set_act_bf16(int mask) { g_act_bf16 = mask; }
static float o_bf16f(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

#define REAL float
#define SFX _f32
#define SQRT sqrtf
#define EXP expf
#define ERF erff
#include "oracle_impl.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef EXP
#undef ERF

#define REAL double
#define SFX _f64
#define SQRT sqrt
#define EXP exp
#define ERF erf
#include "oracle_impl.inc"
