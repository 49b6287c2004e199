/* CPU oracle for the excess-energy band. TEST INFRASTRUCTURE ONLY: the
 * checker for libj2k's device projection, never linked by it.
 *
 * Restates training/jacobi2000_pt/project.py:66-140 (_anomaly_parts,
 * _face_outflow, _excess_terms, excess_energy, project_excess_energy) in
 * f64 throughout: the mathematics, not the f32 rounding of the PyTorch
 * original. Pinned in tests/test_oracle_excess.py by the reference's band
 * tests (training/tests/test_project.py:33-66) and by identities a wrong
 * term breaks: with dt/dx = 0 the excess is energy(x) - g n hbar^2 / 2
 * (sum of eta = 0); the cross term is g/2 sum eta (h_new - h) with h_new
 * the solver's continuity step; scaling (eta, u, v) by s gives
 * s^2 a + s^3 q. Grids are x [3][ny][nx], normalized. */
#include <math.h>
#include <stdlib.h>

#include "j2k_oracle.h"

/* project.py:57-60 */
static double mass_of(const double* x, int n) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i];
    return s / n;
}

/* project.py:78-84: sum over the cells of eta * div(f u) with zero-flux
 * walls, face flux (f_a + f_b)(u_a + u_b)/4. f NULL means f = 1. */
static double eta_outflow(const double* eta, const double* f, const double* u, const double* v, int nx, int ny) {
    double tot = 0.0;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const int i = y * nx + x;
#define F(j) (f ? f[j] : 1.0)
            const double fxr = x + 1 < nx ? (F(i) + F(i + 1)) * (u[i] + u[i + 1]) * 0.25 : 0.0;
            const double fxl = x > 0 ? (F(i - 1) + F(i)) * (u[i - 1] + u[i]) * 0.25 : 0.0;
            const double fyu = y + 1 < ny ? (F(i) + F(i + nx)) * (v[i] + v[i + nx]) * 0.25 : 0.0;
            const double fyd = y > 0 ? (F(i - nx) + F(i)) * (v[i - nx] + v[i]) * 0.25 : 0.0;
#undef F
            tot += eta[i] * (fxr - fxl + fyu - fyd);
        }
    return tot;
}

/* project.py:87-101: a and q of one grid, excess(s) = s^2 a + s^3 q. */
void orc_excess_terms(const double* x, int nx, int ny, const orc_phys* p, double dt_dx, double* a, double* q) {
    const int n = nx * ny;
    const double m = mass_of(x, n);
    double* eta = (double*)malloc(sizeof(double) * (size_t)n);
    double* u = (double*)malloc(sizeof(double) * (size_t)n);
    double* v = (double*)malloc(sizeof(double) * (size_t)n);
    const double hbar = m * p->std[0] + ((double)p->mean[0] + (double)p->depth);
    double se2 = 0.0, ss = 0.0, ses = 0.0;
    for (int i = 0; i < n; i++) {
        eta[i] = (x[i] - m) * p->std[0];
        u[i] = x[n + i] * p->std[1] + p->mean[1];
        v[i] = x[2 * n + i] * p->std[2] + p->mean[2];
        const double s2 = u[i] * u[i] + v[i] * v[i];
        se2 += eta[i] * eta[i];
        ss += s2;
        ses += eta[i] * s2;
    }
    double A = se2 * (0.5 * p->g) + hbar * ss * 0.5;
    double Q = ses * 0.5;
    if (dt_dx != 0.0) {
        const double k = -0.5 * p->g * dt_dx;
        A += hbar * eta_outflow(eta, NULL, u, v, nx, ny) * k;
        Q += eta_outflow(eta, eta, u, v, nx, ny) * k;
    }
    *a = A;
    *q = Q;
    free(eta);
    free(u);
    free(v);
}

double orc_excess_energy(const double* x, int nx, int ny, const orc_phys* p, double dt_dx) {
    double a, q;
    orc_excess_terms(x, nx, ny, p, dt_dx, &a, &q);
    return a + q;
}

/* project.py:112-140, one grid in place; returns the scale applied (1 when
 * the grid is inside the band and left untouched). */
double orc_project_excess_energy(double* x, int nx, int ny, const orc_phys* p, double dt_dx, double prev,
                                 double r_min, double r_max) {
    const int n = nx * ny;
    const double upper = prev * r_max, lower = prev * r_min;
    double a, q;
    orc_excess_terms(x, nx, ny, p, dt_dx, &a, &q);
    const double e = a + q;
    if (!(e > upper || e < lower)) return 1.0;
    const double target = fmin(fmax(e, lower * (1.0 + 1e-5)), upper * (1.0 - 1e-5));
    double s = sqrt(fmax(target / fmax(e, 1e-30), 0.0));
    for (int it = 0; it < 2; it++) {
        const double f = s * s * a + s * s * s * q - target;
        const double df = s * a * 2.0 + s * s * q * 3.0;
        s = s - f / fmax(df, 1e-30);
    }
    s = fmin(fmax(s, 0.0), 2.0);
    const double m = mass_of(x, n);
    for (int i = 0; i < n; i++) {
        x[i] = (x[i] - m) * s + m;
        x[n + i] = ((x[n + i] * p->std[1] + p->mean[1]) * s - p->mean[1]) / p->std[1];
        x[2 * n + i] = ((x[2 * n + i] * p->std[2] + p->mean[2]) * s - p->mean[2]) / p->std[2];
    }
    return s;
}
