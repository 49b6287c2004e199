/* CPU oracle for the jacobi2000 engine forward pass.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP path in
 * jacobi2000_amd/ is compared against. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * Two model forms, each in an f32 build (`_f32`, same loop and summation
 * order as the reference CPU forward) and an f64 twin (`_f64`, same loops in
 * double):
 *
 *  - form P ("patch2d"): the reference engine forward,
 *    crates/engine/src/cpu.rs:249-352, restated loop for loop.
 *  - form C ("cell_st"): one token per cell of a [T, ny, nx] space-time grid,
 *    stencil attention periodic in space and clipped in time, iterated to a
 *    fixed point with per-slice mass/energy projection. The reference has no
 *    such model; its blocks are the reference blocks (cpu.rs:287-313) and its
 *    projection extends crates/engine/src/project.rs. The form is defined by
 *    this file and by the GPU library, which must agree with it.
 *
 * Parity pins (see tests/test_oracle.py):
 *  - form P forward: the pt-tiny goldens (forward_stencil, forward_full,
 *    rollout_1/10/50) written by the reference's PyTorch model.
 *  - mask predicate: the truth table of crates/engine/src/mask.rs:38-57.
 *  - projections: the known-answer tests of crates/engine/src/project.rs:64-102
 *    plus closed-form invariants.
 *  - GELU/erf: against double-precision erf.
 *  - form C: f32 vs f64 twin agreement, brute-force dense attention with an
 *    explicit mask on tiny grids, periodicity (a cyclic shift of the input
 *    shifts the output), and the projection invariants.
 */
#ifndef J2K_ORACLE_H
#define J2K_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Form P shape. mask_kind 0 = Full, 1 = Stencil{radius, dilation}. */
typedef struct {
    int nx, ny, patch, channels, dim, layers, heads, head_dim, ff_inner, meta_dim;
    int mask_kind, radius, dilation;
} orc_cfg_p;

/* Form C shape. Stencil: (2*rs+1)^2 periodic in space x (2*rt+1) clipped in time. */
typedef struct {
    int nx, ny, T, channels, dim, layers, heads, ff_inner, meta_dim;
    int rs, rt;
} orc_cfg_c;

/* Normalizer + physics needed by the projections (crates/engine/src/norm.rs). */
typedef struct {
    float mean[3], std[3], depth, g, r_max;
} orc_phys;

/* Number of weight tensors: form P = 5 + 16*layers + 10, form C = 4 + 16*layers + 4.
 * Order is documented in oracle/oracle.py (P_NAMES / C_NAMES). */

void orc_forward_p_f32(const orc_cfg_p* c, const float* const* w, const float* state, const float* meta, float* incr);
void orc_forward_p_f64(const orc_cfg_p* c, const double* const* w, const double* state, const double* meta, double* incr);

/* Form C model output y[m*3 + ch] for grid x[ch*M + m] (M = T*ny*nx) and known[m]. */
void orc_forward_c_f32(const orc_cfg_c* c, const float* const* w, const float* x, const uint8_t* known, const float* meta, float* y);
void orc_forward_c_f64(const orc_cfg_c* c, const double* const* w, const double* x, const uint8_t* known, const double* meta, double* y);

/* Reference projections, crates/engine/src/project.rs:9-54 (f32 sums, as the reference). */
float orc_mass_f32(const float* x, int n);
float orc_energy_f32(const float* x, const orc_phys* p, int n);
void orc_project_mass_f32(float* x, float target, int n);
void orc_project_energy_f32(float* x, float prev_energy, const orc_phys* p, int n);
/* f64 twins. */
double orc_mass_f64(const double* x, int n);
double orc_energy_f64(const double* x, const orc_phys* p, int n);
void orc_project_mass_f64(double* x, double target, int n);
void orc_project_energy_f64(double* x, double prev_energy, const orc_phys* p, int n);

/* Form C: blank fill, per-slice targets, projection and the whole fixed-point
 * iteration. grid is [3, T, ny, nx]. mass_t/eref_t are [T]. */
void orc_fill_blank(const float* grid, const uint8_t* known, int T, int ny, int nx, float* out);
void orc_slice_targets_f64(const float* x0, const orc_phys* p, int T, int n, double* mass_t, double* eref_t);
/* Applies mass then energy projection on the blank cells of every slice, sums in f64. */
void orc_project_c_f32(float* x, const uint8_t* known, const orc_phys* p, int T, int n, const double* mass_t, const double* eref_t);
void orc_project_c_f64(double* x, const uint8_t* known, const orc_phys* p, int T, int n, const double* mass_t, const double* eref_t);
/* Runs until max|dx| over blank cells < tol or max_iter iterations. x_io holds
 * x0 on entry (blank cells already filled) and the final iterate on return.
 * resid_trace[k] = max|dx| after iteration k (may be NULL). Returns iterations run. */
int orc_iterate_c_f32(const orc_cfg_c* c, const float* const* w, const orc_phys* p, float* x_io, const uint8_t* known,
                      const float* meta, int max_iter, float tol, float* resid_trace);
int orc_iterate_c_f64(const orc_cfg_c* c, const double* const* w, const orc_phys* p, double* x_io, const uint8_t* known,
                      const double* meta, int max_iter, double tol, double* resid_trace);
/* The same, also writing every iterate to states [max_iter][3][T][ny][nx] (may be NULL). */
int orc_iterate_c_states_f32(const orc_cfg_c* c, const float* const* w, const orc_phys* p, float* x_io,
                             const uint8_t* known, const float* meta, int max_iter, float tol, float* resid_trace,
                             float* states);
int orc_iterate_c_states_f64(const orc_cfg_c* c, const double* const* w, const orc_phys* p, double* x_io,
                             const uint8_t* known, const double* meta, int max_iter, double tol, double* resid_trace,
                             double* states);
/* max over slices of max(0, (E_t - r_max*Eref_t)/Eref_t), energies in f64. */
double orc_energy_violation_c(const float* x, const orc_phys* p, int T, int n, const double* eref_t);

/* mask.rs:18-30 predicate, for the truth-table test. */
int orc_mask_allowed(int mask_kind, int radius, int dilation, int tokens_x, int q, int k);
/* GELU with the exact erf form, cpu.rs:67-74. */
float orc_gelu_f32(float x);

/* Threads used by the parallel row loops (OpenMP); 0 = library default. */
void orc_set_threads(int n);
/* Mask of the trunk values rounded to bf16 (0, the default: none;
 * ORC_BF16_ALL: every point where the served bf16 form stores bf16). */
#define ORC_BF16_LN1 1
#define ORC_BF16_QKV 2
#define ORC_BF16_ATTN 4
#define ORC_BF16_LN2 8
#define ORC_BF16_GELU 16
#define ORC_BF16_ALL 31
void orc_set_act_bf16(int mask);

/* Shallow-water solver (crates/solvers/src/shallow_water.rs), sw_oracle.c. */
typedef struct {
    int nx, ny;
    float dx, dt, g, depth, nu;
} orc_sw_params;
/* ShallowWater::steps (shallow_water.rs:149-244), in place. */
void orc_sw_steps(const orc_sw_params* p, float* h, float* u, float* v, int steps);
/* State::add_drop (shallow_water.rs:102-111). */
void orc_sw_add_drop(const orc_sw_params* p, float* h, float cx, float cy, float amp, float sigma);
/* One dump-dataset trajectory (dump_dataset.rs:168-193): traj [3][ticks+1][n]
 * (may be NULL); drops are (tick, field) pairs applied after that tick's
 * step in array order, fields [.][n]. */
void orc_sw_trajectory(const orc_sw_params* p, float* h, float* u, float* v, int ticks, float* traj, int n_drops,
                       const int* drop_tick_field, const float* fields);
/* metrics.rs:9-23. */
float orc_correlation(const float* a, const float* b, int n);

/* The other field scores (crates/engine/src/metrics.rs), metrics_oracle.c. */
float orc_peak(const float* f, int n);
float orc_mean(const float* f, int n);
float orc_std_dev(const float* f, int n);
float orc_bias(const float* model, const float* solver, int n);
/* Radial profile into out (at most cap bins); returns the bin count or -1. */
int orc_radial_profile(const float* f, int nx, int ny, float cx, float cy, double* out, int cap);
float orc_xcorr_lag(const double* model, const double* solver, int n, int max_lag);
/* Excess-energy band (training/jacobi2000_pt/project.py:66-140), f64,
 * excess_oracle.c. x is one normalized grid [3][ny][nx]. */
void orc_excess_terms(const double* x, int nx, int ny, const orc_phys* p, double dt_dx, double* a, double* q);
double orc_excess_energy(const double* x, int nx, int ny, const orc_phys* p, double dt_dx);
double orc_project_excess_energy(double* x, int nx, int ny, const orc_phys* p, double dt_dx, double prev,
                                 double r_min, double r_max);
/* norm.rs:27-34 encode and metrics.rs:109-113 decode_h. */
void orc_encode(const float* h, const float* u, const float* v, int n, const float* mean, const float* stdv,
                float depth, float* out);
void orc_decode_h(const float* x, int n, const float* mean, const float* stdv, float* out);

#ifdef __cplusplus
}
#endif
#endif
