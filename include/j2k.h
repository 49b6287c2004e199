/* libj2k: MI355X (gfx950) HIP implementation of the jacobi2000 engine forward
 * pass behind a plain C ABI.
 *
 * Conventions (one rule for every entry point):
 *  - Every function returns an int status (J2K_OK = 0). On failure
 *    j2k_last_error() returns a message for the calling thread.
 *  - Handles are opaque. A j2k_model owns its device, its weights in HBM, one
 *    HIP stream and its scratch; calls on one handle must not overlap (the
 *    reference GpuModel serialises on one wgpu queue the same way,
 *    crates/engine/src/gpu.rs:374-437).
 *  - Host buffers are caller-owned and only read or written during the call.
 *    `_dev` entry points take device pointers and a hipStream_t (as void*,
 *    NULL = the model's own stream) and return without synchronising.
 *  - Grid layout is the reference's: channel planes c*ny*nx + y*nx + x with
 *    channel 0 = h, 1 = u, 2 = v (crates/engine/src/cpu.rs:7-10). A space-time
 *    grid adds the time slice: c*T*ny*nx + t*ny*nx + y*nx + x.
 */
#ifndef J2K_H
#define J2K_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. 1-5 mirror lean_runtime::Error (crates/lean-runtime/src/error.rs:6-13);
 * 6-9 name the conditions on which the reference loader panics
 * (config.rs:71 missing field, weights.rs:39 non-F32, weights.rs:53 missing
 * tensor, shape mismatches). */
#define J2K_OK 0
#define J2K_ERR_NO_ADAPTER 1     /* no HIP device visible                         */
#define J2K_ERR_NO_DEVICE 2      /* device index invalid or context failed        */
#define J2K_ERR_LIMIT 3          /* a device limit the model needs is not met     */
#define J2K_ERR_MAP 4            /* a host<->device copy failed                   */
#define J2K_ERR_POLL 5           /* a kernel launch or synchronisation failed     */
#define J2K_ERR_BAD_CONFIG 6     /* config.json invalid or missing a field        */
#define J2K_ERR_MISSING_TENSOR 7 /* a tensor the config needs is absent           */
#define J2K_ERR_BAD_DTYPE 8      /* a tensor is not F32                           */
#define J2K_ERR_SHAPE 9          /* a tensor or argument has the wrong shape      */
#define J2K_ERR_INVALID 10       /* bad argument (NULL, out of range, wrong form) */
#define J2K_ERR_OOM 11           /* device memory exhausted / over budget         */
#define J2K_PENDING 12           /* server: result not ready yet                  */

/* Arithmetic of the transformer trunk. LayerNorm statistics, softmax, the
 * residual stream, the fixed-point state and the projections stay f32 (sums
 * in f64) in both. */
#define J2K_PREC_F32 0  /* plain form: f32 operands, f32 MFMA, f32 accumulate   */
#define J2K_PREC_BF16 1 /* served form: bf16 weights/activations, f32 accumulate */
/* served form, compensated: bf16 weights; activations kept f32 and fed to the
 * bf16 MFMA as a hi + lo bf16 pair (two MFMAs per tile), f32 accumulate */
#define J2K_PREC_BF16X2 2
/* FP8: the four block products on E4M3 codes (weights per-tensor scaled at
 * load; activations quantized in the GEMM, each token with the scale of its
 * own maximum, measured just before the product), f32 accumulate; everything
 * else f32. */
#define J2K_PREC_FP8 3

#define J2K_FORM_PATCH2D 0 /* the reference forward (crates/engine/src/cpu.rs:249-352) */
#define J2K_FORM_CELL_ST 1 /* per-cell space-time tokens, fixed-point iteration     */

#define J2K_MASK_FULL 0
#define J2K_MASK_STENCIL 1

/* Config as parsed from config.json (crates/engine/src/config.rs:44-100),
 * plus the cell_st keys. head_dim and ff_inner are inferred from tensor
 * shapes as config.rs:72-73 does. */
typedef struct {
    int form, precision;
    int nx, ny, patch, channels, dim, layers, heads, head_dim, ff_inner, meta_dim;
    int tokens_x, tokens_y, tokens, patch_dim;
    int mask_kind, radius, dilation; /* patch2d: token stencil (mask.rs:18-30)      */
    int time_radius;                 /* cell_st: space radius = radius, periodic    */
    float norm_mean[3], norm_std[3], norm_depth;
    float dx, dt, g, depth, nu, energy_r_max;
    int max_iter;
    float tol;
    int stall_stop; /* cell_st: retire an iteration stuck at its noise floor (default 1) */
    /* The excess-energy band training exported (training/jacobi2000_pt/export.py:
     * 34-35,67-68): upper and lower ratio, NaN when config.json has none (or null).
     * The reference engine reads neither; see j2k_project_excess_energy. */
    float excess_r, excess_r_min;
} j2k_config;

/* Thread-local message of the last failure on this thread ("" if none). */
const char* j2k_last_error(void);

/* Number of visible HIP devices (0 with J2K_ERR_NO_ADAPTER when none). */
int j2k_device_count(int* n);

/* Model::from_bytes (crates/engine/src/lib.rs:45-50) + GpuModel::new
 * (crates/engine/src/gpu.rs:212-218): parse config.json, read the F32
 * safetensors bytes, upload every weight to `device`. `precision` is
 * J2K_PREC_F32 or J2K_PREC_BF16. */
int j2k_model_create(const uint8_t* safetensors, size_t n_bytes, const char* config_json, int device, int precision,
                     void** out_model);
/* GpuModel's Drop (gpu.rs:199-209): frees every device buffer of the model. */
int j2k_model_destroy(void* model);
/* GpuModel::cfg (gpu.rs:427-429). */
int j2k_model_config(const void* model, j2k_config* out);

/* Autotuner, no reference counterpart (the reference has one fixed launch
 * shape per kernel): times the candidate launch shapes of the BF16 / BF16X2
 * trunk (GEMM block height, attention K/V staging: joint, split, per time
 * plane or two queries per lane; the d = 128 token-local work apart or fused
 * into one launch of four waves (in a row or 2 x 2, or with 64-column MLP
 * chunks) or of eight waves; the d = 256 token-local work as tile GEMMs and
 * LayerNorm launches, with the MLP in one launch, or all in one launch) on
 * synthetic packs of 2^b tokens for
 * every b from 12 up to that of max_tokens (at most 2^20) not yet tuned, and
 * keeps the fastest per bucket of packed tokens. A choice never changes a
 * result. j2k_iterate* and j2k_server_create run it on first use unless the
 * environment sets J2K_AUTOTUNE=0; J2K_TUNE_FORCE =
 * "gemm_bm=64|128,attn_kv_mode=0..4,rows_fused=0..4" instead fixes the choices
 * of models created after it is set. max_tokens = 0 only reports. `report`
 * (may be NULL) receives the table as JSON text, cut to report_len bytes.
 * F32 models keep their fixed shapes. Synchronises the model's stream; not
 * concurrently with other calls on the model. */
int j2k_model_autotune(void* model, long long max_tokens, char* report, size_t report_len);

/* Profiling hook, no reference counterpart (the reference timed passes with
 * GPU timestamp queries it did not commit): enable = 1 / 0 turns HIP-event
 * timing of the model's launches on or off and clears the counters (-1 keeps
 * the state). `out` (may be NULL) receives 21 doubles, per kernel class
 * (GEMM, attention, LayerNorm, projection, row-panel kernel making the next
 * QKV, other row-panel launches, the first block's row-panel launch made
 * from the embedding): total ms, total algorithmic work (FLOPs for GEMM,
 * bytes otherwise), launches. Synchronises the model's stream. */
int j2k_model_timing(void* model, int enable, double* out);

/* ---- patch2d form (the reference forward) ----------------------------- */

/* Model::forward / GpuModel::forward (lib.rs:53-55, gpu.rs:374-425):
 * state [channels, ny, nx] normalized, meta [meta_dim] -> increment, same shape. */
int j2k_forward(void* model, const float* state, const float* meta, float* incr);
/* Model::step / GpuModel::step (lib.rs:58-60, cpu.rs:84-88, gpu.rs:431-437): state + forward. */
int j2k_step(void* model, const float* state, const float* meta, float* next);
/* Batched forward on device: `batch` grids [batch, channels, ny, nx], one meta
 * vector for all; out = forward(state) (+ state when add_state). No reference
 * counterpart (the reference runs batch 1, lib.rs:4); it is the batched form of
 * j2k_forward that rollout windows and serving use. */
int j2k_forward_batch_dev(void* model, int batch, const float* d_state, const float* d_meta, float* d_out,
                          int add_state, void* stream);
/* The rollout loop of crates/engine-cli/src/main.rs:117-121 and the projected
 * loop of eval.rs:144-156, kept on the device: `steps` times
 *   x <- step(x); if proj_mass: project_mass(x, mass(x_0));
 *   if proj_energy: project_energy(x, energy(x_prev)) (mass first, eval.rs order)
 * with the projection sums in f64 on the device. energy_trace (may be NULL)
 * receives steps+1 energies (f64, physical units); max_violation (may be NULL)
 * receives max_t max(0, (E_t - r_max E_{t-1}) / E_{t-1}) (eval.rs:156). */
int j2k_rollout(void* model, float* state_io, const float* meta, int steps, int proj_mass, int proj_energy,
                double* energy_trace, double* max_violation);
/* The same loop on device buffers; energy_trace is a device buffer of steps+1 doubles or NULL. */
int j2k_rollout_dev(void* model, float* d_state_io, const float* d_meta, int steps, int proj_mass, int proj_energy,
                    double* d_energy_trace, void* stream);
/* Jacobi-over-time window: window_io holds T grids [T, channels, ny, nx];
 * slice 0 is the initial state. Each iteration replaces slices 1..T-1 by
 * project(step(slice t-1 of the previous iterate)) in one batched forward;
 * after k iterations slices 0..k equal a k-step rollout exactly (the fixed
 * point of the rollout, main.rs:117-121, reached after T-1 iterations). */
int j2k_iterate_window(void* model, float* window_io, const float* meta, int T, int iters, int proj_mass,
                       int proj_energy);

/* project::mass / project_mass / energy / project_energy (project.rs:9-54) on
 * one host grid [3, n], sums in f64 on the device. */
int j2k_mass(void* model, const float* x, int n, double* out);
int j2k_energy(void* model, const float* x, int n, double* out);
int j2k_project_mass(void* model, float* x, int n, float target);
int j2k_project_energy(void* model, float* x, int n, float prev_energy);

/* ---- cell_st form (per-cell space-time fixed point) ------------------- */

/* Blank fill used before iterating (host only): a blank cell takes the same
 * (y, x) cell of the nearest fully known slice (ties to the earlier slice),
 * else the mean of its slice's known cells in that channel, else the mean of
 * all known cells, else 0. grid [3,T,ny,nx], known [T,ny,nx] (0/1). */
int j2k_fill_blank(const float* grid, const uint8_t* known, int T, int ny, int nx, float* out);

/* Runs the fixed point x_{k+1} = P(clamp_known(x_0 + f(x_k))) from x_0 =
 * grid_io (blank cells already filled) until max|x_{k+1} - x_k| over blank
 * cells < tol or max_iter iterations (tol < 0 / max_iter <= 0: config values),
 * or, with the config's stall_stop, when the residual stops contracting (two
 * consecutive iterations not below 0.99x the best so far).
 * P projects every slice on its blank cells: mass to the mass of x_0's slice,
 * then energy to r_max times x_0's slice energy, sums in f64 on the device.
 * Writes the final iterate to grid_io; iters, resid, violation may be NULL. */
int j2k_iterate(void* model, float* grid_io, const uint8_t* known, const float* meta, int T, int ny, int nx,
                int max_iter, float tol, int* iters, float* resid, double* violation);
/* Device-pointer form; synchronises for the per-iteration convergence test
 * and once at the end (the violation; the grid is written on return).
 * d_grid [3,T,ny,nx], d_known [T,ny,nx] u8, d_meta [meta_dim].
 * From the second call with the same grid shape and the same d_grid, d_known
 * and d_meta pointers on, each iteration (the first with the request's
 * set-up, the unpack and the residual copy) replays from a HIP graph
 * (captured on that call; results bit-identical to the eager launches)
 * unless per-kernel timing is on or the environment held J2K_GRAPHS=0 when
 * the model was created. j2k_iterate uploads into the model's own buffers,
 * so it replays from its second call with a grid shape. */
int j2k_iterate_dev(void* model, float* d_grid, const uint8_t* d_known, const float* d_meta, int T, int ny, int nx,
                    int max_iter, float tol, int* iters, float* resid, double* violation, void* stream);
/* One model evaluation y = f(x) of the cell_st form: y [T*ny*nx, 3]. */
int j2k_cell_forward_dev(void* model, const float* d_grid, const uint8_t* d_known, const float* d_meta, int T, int ny,
                         int nx, float* d_y, void* stream);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the same
 * evaluation as j2k_cell_forward_dev, bit for bit, with every scratch buffer
 * (tables, packed features, residual stream, activations) carved from a
 * fixed device region of J2K_NOALLOC_MAX_BYTES that the library holds on each
 * device it runs on until the process ends; it calls no hipMalloc or
 * hipFree. A request whose scratch does not fit is refused with
 * J2K_ERR_OOM before any device work (about 43,000 tokens in BF16X2 and
 * 26,000 in F32 at d = 128, ff 512). Every call on one device shares the
 * region, so calls that overlap on one device race; the call uploads the
 * grid's tables with synchronous copies (not for use under stream capture). */
#define J2K_NOALLOC_MAX_BYTES ((size_t)128 << 20)
int j2k_cell_forward_dev_noalloc(void* model, const float* d_grid, const uint8_t* d_known, const float* d_meta, int T,
                                 int ny, int nx, float* d_y, void* stream);

/* ---- serving (continuous batching on one GPU) -------------------------- */

typedef struct {
    uint64_t id;
    int iters;          /* fixed-point iterations run                            */
    int converged;      /* 1 if max|dx| < tol                                    */
    int stalled;        /* 1 if retired at its noise floor before tol / max_iter */
    float resid;        /* last max|dx|                                          */
    double violation;   /* max relative energy violation over slices             */
    double submit_s, start_s, done_s; /* monotonic seconds: submit, admit, retire */
} j2k_result;

/* A persistent request loop on the model's device. Requests are admitted
 * while their device state fits `hbm_budget_bytes` and at most `max_active`
 * are on the GPU; all active requests' tokens are packed into one ragged
 * launch per layer; a request retires on its own convergence, and its
 * result lands on the host while the next iteration runs (or at once when
 * there is none): poll reports it from then on. */
int j2k_server_create(void* model, size_t hbm_budget_bytes, int max_active, void** out_server);
/* Queues a request (copies its inputs). grid [3,T,ny,nx] with blank cells
 * filled or not (the server fills them), known [T,ny,nx], meta [meta_dim]. */
int j2k_server_submit(void* server, const float* grid, const uint8_t* known, const float* meta, int T, int ny, int nx,
                      uint64_t* out_id);
/* Queues a chained rollout, the served form of the reference's rollout loop
 * (crates/engine-cli/src/main.rs:117-127) run window-parallel in time:
 * state [3,ny,nx] is tick 0; windows of `window` slices overlap by one
 * slice (each window's last slice is the next one's known slice 0, the
 * others start as copies of it) until `ticks` slices are covered. The result
 * (poll's grid_out) is the trajectory [3,ticks,ny,nx]; its iters sum over
 * windows, violation is the worst window's, converged means every window
 * converged. The budget is charged one window. */
int j2k_server_submit_rollout(void* server, const float* state, const float* meta, int ny, int nx, int window,
                              int ticks, uint64_t* out_id);
/* Per-window figures of a retired rollout: *count windows; for the first n,
 * latency (seconds from the previous window's retirement, or from submit),
 * relative energy violation, iterations. J2K_PENDING before retirement. */
int j2k_server_windows(void* server, uint64_t id, int n, int* count, double* latency_s, double* violation,
                       int* iters);
/* Non-blocking: J2K_OK and the result (and the grid into grid_out if not
 * NULL) once retired, J2K_PENDING before. */
int j2k_server_poll(void* server, uint64_t id, j2k_result* out, float* grid_out);
/* Blocks until every submitted request has retired. */
int j2k_server_drain(void* server);
/* Counters: iterations launched, graph replays, peak active requests, peak bytes. */
int j2k_server_stats(void* server, uint64_t* launches, uint64_t* graph_replays, int* peak_active, size_t* peak_bytes);
int j2k_server_destroy(void* server);

/* ---- scores and the rollout's truth partner -------------------------------- */

/* eval-metrics' scores of one trajectory at one horizon
 * (crates/engine-cli/src/eval.rs:158-172). */
typedef struct {
    float corr;            /* metrics::correlation over (h, u, v), normalized fields */
    float corr_h;          /* the same over h                                       */
    float amplitude_ratio; /* peak |h| of the model over the solver's, decoded h    */
    float lag;             /* wavefront lag in cells: xcorr_lag of the radial |h|   */
                           /* profiles about the drop centre, max lag given         */
    float bias;            /* (mean model h - mean solver h) / std solver h         */
} j2k_eval_score;

/* Scores of `batch` patch2d trajectories on the device: d_x [batch][3][ny][nx]
 * normalized model states, d_h / d_u / d_v [batch][ny][nx] the solver's
 * physical states (encoded with the model's normalizer, norm.rs:27-34),
 * centers [batch][2] the drop centres (cx, cy) in cells (host), max_lag the
 * profile lag range (8 in eval.rs:167). Sums run in f64 in a fixed order;
 * the lag is finished on the host as metrics.rs:78-107 does. out [batch]
 * (host). Synchronises the stream (NULL: the model's). */
int j2k_score_dev(void* model, int batch, const float* d_x, const float* d_h, const float* d_u, const float* d_v,
                  const float* centers, int max_lag, j2k_eval_score* out, void* stream);
/* The same from host buffers. */
int j2k_score(void* model, int batch, const float* x, const float* h, const float* u, const float* v,
              const float* centers, int max_lag, j2k_eval_score* out);
/* metrics::correlation (metrics.rs:9-23) of d_a*sa + ta against d_b*sb + tb
 * over n values, on the current device; the f32 result goes to d_out[0]
 * (device). decode_h is sa = std[0], ta = mean[0]; the solver's h - depth
 * is sb = 1, tb = -depth. */
int j2k_correlation_dev(const float* d_a, float sa, float ta, const float* d_b, float sb, float tb, int n,
                        float* d_out, void* stream);
/* The rollout CLI's loop (crates/engine-cli/src/main.rs:116-127) with its
 * truth partner on the device too: the solver (a j2k_sw handle on the
 * model's grid and device) advances the physical state h, u, v [ny][nx] one
 * tick per step, the model steps the normalized state x [3][ny][nx] with the
 * mass projection to the initial mass (f64 sums), and corr[step-1] gets the
 * correlation of decode_h(x) with the solver's h - depth. Everything stays
 * on the device until the end: x and h, u, v return their final values,
 * corr [steps] the trace (host buffers). */
int j2k_rollout_truth(void* model, void* solver, float* x_io, const float* meta, float* h, float* u, float* v,
                      int steps, float* corr);
/* The same on device buffers (corr: device, [steps]); enqueued on `stream`
 * (NULL: the model's). */
int j2k_rollout_truth_dev(void* model, void* solver, float* d_x, const float* d_meta, float* d_h, float* d_u,
                          float* d_v, int steps, float* d_corr, void* stream);

/* ---- excess-energy band (training/jacobi2000_pt/project.py:66-140) -------
 *
 * The band projection of training, on the device, for patch2d grids with the
 * model's normalizer and g: x [batch][3][ny][nx] normalized. Per-cell values
 * are f32 as the reference's tensor ops compute them, the sums f64 (the
 * reference sums in f32). dt_dx = dt/dx adds the forward-backward solver's
 * cross term; 0 gives the plain excess energy. */

/* project.py:98-108 excess_energy of each grid into out [batch] (host). */
int j2k_excess_energy(void* model, int batch, const float* x, float dt_dx, double* out);
int j2k_excess_energy_dev(void* model, int batch, const float* d_x, float dt_dx, double* d_out, void* stream);
/* project.py:111-140 project_excess_energy in place: grids whose excess lies
 * outside [prev*r_min, prev*r_max] get (eta, u, v) scaled by one s (mass
 * exact), the others are untouched; r_min = 0 is the upper-only form. prev
 * [batch] is each grid's excess before the step. */
int j2k_project_excess_energy(void* model, int batch, float* x, const double* prev, float dt_dx, float r_min,
                              float r_max);
int j2k_project_excess_energy_dev(void* model, int batch, float* d_x, const double* d_prev, float dt_dx, float r_min,
                                  float r_max, void* stream);

/* ---- shallow-water solver (crates/solvers/src/shallow_water.rs) ----------
 *
 * The reference solver stepped on the device, bit for bit: the same f32
 * operations in the same order (no contraction) as ShallowWater::step
 * (shallow_water.rs:149-238). A state is three fields h, u, v of ny*nx
 * floats, row-major (index y*nx + x); a batch of B states holds each field
 * as [B][ny][nx]. */

/* Params (shallow_water.rs:16-29). */
typedef struct {
    int nx, ny;
    float dx, dt, g, depth, nu;
} j2k_sw_params;

/* ShallowWater::new (shallow_water.rs:136-147): a solver for one grid on
 * `device`. nx, ny >= 1 and nx*ny < 2^31 (J2K_ERR_INVALID otherwise). */
int j2k_sw_create(const j2k_sw_params* params, int device, void** out_solver);
int j2k_sw_destroy(void* solver);
/* ShallowWater::steps (shallow_water.rs:240-244) on `batch` states in host
 * buffers, in place. */
int j2k_sw_steps(void* solver, float* h, float* u, float* v, int batch, int steps);
/* The same on device buffers, enqueued on `stream` (NULL: the solver's own). */
int j2k_sw_steps_dev(void* solver, float* d_h, float* d_u, float* d_v, int batch, int steps, void* stream);

/* A drop of dump-dataset's schedule (dump_dataset.rs:85-95, apply_pending):
 * after step `tick` (1-based) of its state, h += fields[field][ny][nx], the
 * bump State::add_drop adds (see j2k_sw_drop_field). */
typedef struct {
    int tick, field;
} j2k_sw_drop;

/* Solver trajectories, dump-dataset's loop (dump_dataset.rs:168-193) over
 * `batch` independent states: state b runs min(ticks[b], max_ticks) steps;
 * frame 0 is its input, frame t the state after step t and after the drops
 * of tick t (applied in array order). Frames of state b go to
 * traj + offsets[b], laid out [3][frames][ny][nx] (every h frame, then u,
 * then v: the traj-NNNNN.f32 layout) with frames = min(ticks[b], max_ticks)+1;
 * traj NULL records nothing. Drops of state b are drops[drop_begin[b] ..
 * drop_begin[b+1]); drop_begin NULL means none. ticks, offsets, drop_begin,
 * drops and fields are device arrays; h, u, v advance in place. */
int j2k_sw_trajectories_dev(void* solver, float* d_h, float* d_u, float* d_v, int batch, const int* d_ticks,
                            int max_ticks, const int64_t* d_offsets, float* d_traj, const int* d_drop_begin,
                            const j2k_sw_drop* d_drops, const float* d_fields, void* stream);
/* The same from host buffers: ticks[batch]; traj holds every state's frames
 * back to back (state b at the sum over earlier states of 3*(ticks+1)*ny*nx);
 * drops given as n_drops (state, tick, field) triples, applied per state in
 * array order; fields [n_fields][ny][nx]. */
int j2k_sw_trajectories(void* solver, float* h, float* u, float* v, int batch, const int* ticks, float* traj,
                        int n_drops, const int* drop_state_tick_field, const float* fields, int n_fields);
/* State::add_drop's bump (shallow_water.rs:102-111) on the host, as the
 * reference computes it in f32: out[y*nx+x] = amp*exp(-(dx*dx+dy*dy)*inv),
 * inv = 1/(2*sigma*sigma). Adding it to h is State::add_drop. */
int j2k_sw_drop_field(const j2k_sw_params* params, float cx, float cy, float amp, float sigma, float* out);
/* Kernel choice, no reference counterpart: 0 picks by grid size (a grid of
 * up to 6,826 cells runs every step of a state inside one workgroup's
 * LDS, a larger one runs one launch per step through HBM), 1 forces the
 * resident kernel (J2K_ERR_LIMIT if the grid does not fit), 2 forces the
 * HBM one. Both give identical results. */
int j2k_sw_set_kernel(void* solver, int kernel);

#ifdef __cplusplus
}
#endif
#endif
