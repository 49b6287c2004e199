/* CPU oracle for the shallow-water solver. TEST INFRASTRUCTURE ONLY: the
 * checker for libj2k's device solver, never linked by it.
 *
 * Restates crates/solvers/src/shallow_water.rs loop for loop (face flux
 * arrays, h_new, then momentum on the new surface) and the trajectory loop
 * of crates/engine-cli/src/dump_dataset.rs:168-193. Built with
 * -ffp-contract=off: every a*b+c rounds twice, as Rust's f32 does.
 *
 * Pinned in tests/test_oracle_solver.py by the reference's own tests
 * (shallow_water.rs:260-322: flat stays flat, mass conserved, mirror
 * symmetry, bounded long run) and by closed forms a wrong index, sign or
 * factor breaks: the linear gravity-wave frequency of the forward-backward
 * scheme, one-step viscous decay of a cosine shear mode, one-step advection
 * of linear velocity profiles, and the wall fluxes of a uniform flow. */
#include "j2k_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* shallow_water.rs:115-124 */
static int mirror(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

/* shallow_water.rs:149-238, one step, in place. */
static void sw_step(const orc_sw_params* p, float* h, float* u, float* v, float* fx, float* fy, float* hn, float* un,
                    float* vn) {
    const int nx = p->nx, ny = p->ny;
#define IDX(x, y) ((size_t)(y) * nx + (x))
    for (int y = 0; y < ny; y++) {
        for (int x = 1; x < nx; x++) {
            const size_t a = IDX(x - 1, y), b = IDX(x, y);
            fx[(size_t)y * (nx + 1) + x] = 0.5f * (h[a] + h[b]) * 0.5f * (u[a] + u[b]);
        }
        fx[(size_t)y * (nx + 1)] = 0.0f;
        fx[(size_t)y * (nx + 1) + nx] = 0.0f;
    }
    for (int x = 0; x < nx; x++) {
        for (int y = 1; y < ny; y++) {
            const size_t a = IDX(x, y - 1), b = IDX(x, y);
            fy[(size_t)y * nx + x] = 0.5f * (h[a] + h[b]) * 0.5f * (v[a] + v[b]);
        }
        fy[x] = 0.0f;
        fy[(size_t)ny * nx + x] = 0.0f;
    }
    const float r = p->dt / p->dx;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const float dfx = fx[(size_t)y * (nx + 1) + x + 1] - fx[(size_t)y * (nx + 1) + x];
            const float dfy = fy[(size_t)(y + 1) * nx + x] - fy[(size_t)y * nx + x];
            hn[IDX(x, y)] = h[IDX(x, y)] - r * (dfx + dfy);
        }
    const float inv2dx = 1.0f / (2.0f * p->dx);
    const float invdx2 = 1.0f / (p->dx * p->dx);
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const size_t i = IDX(x, y);
            const float sxm = x == 0 ? -1.0f : 1.0f, sxp = x + 1 == nx ? -1.0f : 1.0f;
            const float sym = y == 0 ? -1.0f : 1.0f, syp = y + 1 == ny ? -1.0f : 1.0f;
            const float hl = hn[IDX(mirror(x - 1, nx), y)], hr = hn[IDX(mirror(x + 1, nx), y)];
            const float hd = hn[IDX(x, mirror(y - 1, ny))], hu = hn[IDX(x, mirror(y + 1, ny))];
            const float ul = sxm * u[IDX(mirror(x - 1, nx), y)], ur = sxp * u[IDX(mirror(x + 1, nx), y)];
            const float ud = u[IDX(x, mirror(y - 1, ny))], uu = u[IDX(x, mirror(y + 1, ny))];
            const float vl = v[IDX(mirror(x - 1, nx), y)], vr = v[IDX(mirror(x + 1, nx), y)];
            const float vd = sym * v[IDX(x, mirror(y - 1, ny))], vu = syp * v[IDX(x, mirror(y + 1, ny))];
            const float u0 = u[i], v0 = v[i];
            const float dudx = (ur - ul) * inv2dx, dudy = (uu - ud) * inv2dx;
            const float dvdx = (vr - vl) * inv2dx, dvdy = (vu - vd) * inv2dx;
            const float dhdx = (hr - hl) * inv2dx, dhdy = (hu - hd) * inv2dx;
            const float lap_u = (ul + ur + ud + uu - 4.0f * u0) * invdx2;
            const float lap_v = (vl + vr + vd + vu - 4.0f * v0) * invdx2;
            un[i] = u0 + p->dt * (-u0 * dudx - v0 * dudy - p->g * dhdx + p->nu * lap_u);
            vn[i] = v0 + p->dt * (-u0 * dvdx - v0 * dvdy - p->g * dhdy + p->nu * lap_v);
        }
#undef IDX
    const size_t n = (size_t)nx * ny;
    memcpy(h, hn, n * sizeof(float));
    memcpy(u, un, n * sizeof(float));
    memcpy(v, vn, n * sizeof(float));
}

typedef struct {
    float *fx, *fy, *hn, *un, *vn;
} sw_scratch;

static sw_scratch scratch_new(const orc_sw_params* p) {
    const size_t n = (size_t)p->nx * p->ny;
    sw_scratch s;
    s.fx = (float*)malloc(sizeof(float) * (size_t)(p->nx + 1) * p->ny);
    s.fy = (float*)malloc(sizeof(float) * (size_t)p->nx * (p->ny + 1));
    s.hn = (float*)malloc(sizeof(float) * n);
    s.un = (float*)malloc(sizeof(float) * n);
    s.vn = (float*)malloc(sizeof(float) * n);
    return s;
}

static void scratch_free(sw_scratch* s) {
    free(s->fx);
    free(s->fy);
    free(s->hn);
    free(s->un);
    free(s->vn);
}

void orc_sw_steps(const orc_sw_params* p, float* h, float* u, float* v, int steps) {
    sw_scratch s = scratch_new(p);
    for (int t = 0; t < steps; t++) sw_step(p, h, u, v, s.fx, s.fy, s.hn, s.un, s.vn);
    scratch_free(&s);
}

/* shallow_water.rs:102-111 */
void orc_sw_add_drop(const orc_sw_params* p, float* h, float cx, float cy, float amp, float sigma) {
    const float inv = 1.0f / (2.0f * sigma * sigma);
    for (int y = 0; y < p->ny; y++)
        for (int x = 0; x < p->nx; x++) {
            const float dx = (float)x - cx, dy = (float)y - cy;
            h[(size_t)y * p->nx + x] += amp * expf(-(dx * dx + dy * dy) * inv);
        }
}

/* dump_dataset.rs:168-193 for one state: frame 0, then per tick a step, the
 * drops of that tick in the given order (apply_pending, :85-95), a frame.
 * traj is [3][ticks+1][n]; drops are (tick, field) pairs into fields[.][n]. */
void orc_sw_trajectory(const orc_sw_params* p, float* h, float* u, float* v, int ticks, float* traj, int n_drops,
                       const int* drop_tick_field, const float* fields) {
    const size_t n = (size_t)p->nx * p->ny, plane = (size_t)(ticks + 1) * n;
    sw_scratch s = scratch_new(p);
    for (int t = 0; t <= ticks; t++) {
        if (t > 0) {
            sw_step(p, h, u, v, s.fx, s.fy, s.hn, s.un, s.vn);
            for (int d = 0; d < n_drops; d++)
                if (drop_tick_field[2 * d] == t) {
                    const float* f = fields + (size_t)drop_tick_field[2 * d + 1] * n;
                    for (size_t i = 0; i < n; i++) h[i] += f[i];
                }
        }
        if (traj) {
            memcpy(traj + (size_t)t * n, h, n * sizeof(float));
            memcpy(traj + plane + (size_t)t * n, u, n * sizeof(float));
            memcpy(traj + 2 * plane + (size_t)t * n, v, n * sizeof(float));
        }
    }
    scratch_free(&s);
}

/* crates/engine/src/metrics.rs:9-23: Pearson correlation in f64, 0 when
 * either field is constant. */
float orc_correlation(const float* a, const float* b, int n) {
    double ma = 0, mb = 0;
    for (int i = 0; i < n; i++) {
        ma += a[i];
        mb += b[i];
    }
    ma /= n;
    mb /= n;
    double sab = 0, saa = 0, sbb = 0;
    for (int i = 0; i < n; i++) {
        const double dx = a[i] - ma, dy = b[i] - mb;
        sab += dx * dy;
        saa += dx * dx;
        sbb += dy * dy;
    }
    if (saa <= 0 || sbb <= 0) return 0.0f;
    return (float)(sab / sqrt(saa * sbb));
}
