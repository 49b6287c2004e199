// The C ABI of libj2k (include/j2k.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "abi.h"
#include "cell.h"
#include "model.h"
#include "score.h"

namespace j2k {

thread_local std::string g_err;
void set_last_error(const std::string& msg) { g_err = msg; }

namespace {

Model& model_of(void* h) {
    if (!h) fail(J2K_ERR_INVALID, "model handle is NULL");
    return *static_cast<Model*>(h);
}

void need(const void* p, const char* what) {
    if (!p) fail(J2K_ERR_INVALID, std::string(what) + " is NULL");
}

void need_form(const Model& m, int form, const char* fn) {
    if (m.cfg.form != form)
        fail(J2K_ERR_INVALID, std::string(fn) + ": wrong model form (" +
                                  (form == J2K_FORM_PATCH2D ? "needs patch2d" : "needs cell_st") + ")");
}

hipStream_t stream_of(Model& m, void* s) { return s ? static_cast<hipStream_t>(s) : m.stream; }

void h2d(void* d, const void* h, size_t bytes, hipStream_t s) {
    hip_check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s), "host->device copy", J2K_ERR_MAP);
}
void d2h(void* h, const void* d, size_t bytes, hipStream_t s) {
    hip_check(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s), "device->host copy", J2K_ERR_MAP);
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize", J2K_ERR_POLL);
}
void sync(hipStream_t s) { hip_check(hipStreamSynchronize(s), "hipStreamSynchronize", J2K_ERR_POLL); }

// ---- patch2d rollout ----------------------------------------------------------

void rollout_dev(Model& m, hipStream_t s, float* x, const float* meta, int steps, int pm, int pe, double* etrace) {
    const j2k_config& c = m.cfg;
    const int n = c.ny * c.nx;
    float* xb = m.get<float>(S_WIN_B, 3 * (size_t)n);
    double* sums = m.get<double>(S_SUMS, 2 + (etrace ? 0 : (size_t)steps + 1));
    double* mass0 = sums;
    double* e = etrace ? etrace : sums + 2;
    grid_sums(s, 1, n, x, mass0, e, c.g, c.norm_mean, c.norm_std, c.norm_depth);
    float* cur = x;
    float* nxt = xb;
    for (int t = 1; t <= steps; t++) {
        forward_p(m, s, 1, cur, meta, nxt, true);
        project_p(s, 1, n, nxt, mass0, pm, e + t - 1, pe, e + t, c.g, c.norm_mean, c.norm_std, c.norm_depth,
                  c.energy_r_max);
        std::swap(cur, nxt);
    }
    if (cur != x) HIPCHK(hipMemcpyAsync(x, cur, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, s));
}

// main.rs:116-127 with the solver's frames: the truth partner runs every
// tick in one launch, then per step the model step, the mass projection and
// the correlation of decode_h(x) against frame t's h - depth.
void rollout_truth_dev(Model& m, hipStream_t s, void* solver, float* x, const float* meta, float* h, float* u,
                       float* v, int steps, float* corr) {
    const j2k_config& c = m.cfg;
    int snx = 0, sny = 0, sdev = 0;
    float sdepth = 0.0f;
    sw_grid_of(solver, &snx, &sny, &sdepth, &sdev);
    if (snx != c.nx || sny != c.ny)
        fail(J2K_ERR_SHAPE, "rollout truth: the solver's grid is " + std::to_string(snx) + "x" + std::to_string(sny) +
                                ", the model's " + std::to_string(c.nx) + "x" + std::to_string(c.ny));
    if (sdev != m.device) fail(J2K_ERR_INVALID, "rollout truth: the solver and the model are on different devices");
    const int n = c.ny * c.nx;
    const float* frames = sw_frames_on(solver, s, h, u, v, steps);
    float* xb = m.get<float>(S_WIN_B, 3 * (size_t)n);
    double* mass0 = m.get<double>(S_SUMS, 2);
    grid_sums(s, 1, n, x, mass0, mass0 + 1, c.g, c.norm_mean, c.norm_std, c.norm_depth);
    float* cur = x;
    float* nxt = xb;
    for (int t = 1; t <= steps; t++) {
        forward_p(m, s, 1, cur, meta, nxt, true);
        project_p(s, 1, n, nxt, mass0, 1, nullptr, 0, nullptr, c.g, c.norm_mean, c.norm_std, c.norm_depth,
                  c.energy_r_max);
        std::swap(cur, nxt);
        // metrics.rs:111-113 decode_h against main.rs:123 h - depth
        corr_affine(s, cur, c.norm_std[0], c.norm_mean[0], frames + (size_t)t * n, 1.0f, -sdepth, n, corr + t - 1);
    }
    if (cur != x) HIPCHK(hipMemcpyAsync(x, cur, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, s));
}

double violation_of(const std::vector<double>& e, float r_max) {
    double worst = 0;
    for (size_t t = 1; t < e.size(); t++) worst = std::max(worst, std::max(0.0, (e[t] - e[t - 1] * r_max) / e[t - 1]));
    return worst;
}

// ---- cell_st single request ---------------------------------------------------

void iterate_dev(Model& m, hipStream_t s, float* d_grid, const uint8_t* d_known, const float* d_meta, int T, int ny,
                 int nx, int max_iter, float tol, int* iters, float* resid, double* violation) {
    need_form(m, J2K_FORM_CELL_ST, "j2k_iterate");
    if (max_iter <= 0) max_iter = m.cfg.max_iter;
    if (tol < 0) tol = m.cfg.tol;
    if (m.tune.enabled) autotune(m, (long)T * ny * nx);  // before the tables: it reuses their slots
    CellRun run;
    cell_setup(m, s, CellDims{T, ny, nx}, nullptr, run);
    unsigned long long* vb = reinterpret_cast<unsigned long long*>(run.pack.resid + 2);
    run.pack.viol = vb;  // the projection leaves the violation of each iterate
    const CellPack& p = run.pack;
    float* x0 = m.get<float>(S_X0, 4 * (size_t)p.M);
    float* xa = m.get<float>(S_XA, 4 * (size_t)p.M);
    float* xb = m.get<float>(S_XB, 4 * (size_t)p.M);
    // the request's meta vector, packed start and projection targets
    auto prologue = [&] {
        cell_pack_targets(s, p.n_slices, p.slices, p.M, d_grid, d_known, x0, xa, p.mass_t, p.eref_t, m.cfg.g,
                          m.cfg.norm_mean, m.cfg.norm_std, m.cfg.norm_depth, m.cfg.meta_dim, m.cfg.dim, d_meta,
                          m.meta_w, m.meta_b, p.mvec);
    };
    int it = 0;
    float r = 0.f;
    StallWatch watch;
    // Replay the iterations from graphs once this grid and these buffers
    // have run eagerly (not while per-kernel timing records events).
    auto& cg = m.cell_graph;
    if (cg.T != T || cg.ny != ny || cg.nx != nx || cg.x0 != x0 || cg.xa != xa || cg.xb != xb || cg.grid != d_grid ||
        cg.known != d_known || cg.meta != d_meta || cg.epoch != m.graph_epoch) {
        m.drop_cell_graphs();
        cg.T = T, cg.ny = ny, cg.nx = nx, cg.x0 = x0, cg.xa = xa, cg.xb = xb, cg.epoch = m.graph_epoch;
        cg.grid = d_grid, cg.known = d_known, cg.meta = d_meta;
    }
    if (cg.calls < 2) cg.calls++;
    const bool replay = !m.timer.on && !m.graphs_off && cg.calls >= 2;
    // One iteration (the first also sets the request up), then the unpack of
    // its iterate into d_grid and the copy of the residual word and the
    // violation after it (the projection leaves both) to h_words, so a
    // request that stops there needs no second round trip; a later
    // iteration overwrites both.
    auto iteration = [&](int kind, float* xc, float* xn) {
        auto body = [&] {
            if (kind == 0) prologue();
            cell_iteration(m, s, p, x0, xc, xn, d_grid);  // the projection also unpacks into d_grid
            hip_check(hipMemcpyAsync(m.h_words, p.resid, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, s),
                      "device->host copy", J2K_ERR_MAP);
        };
        if (replay && !cg.g[kind]) {
            hipGraph_t g = nullptr;
            const unsigned before = m.graph_epoch;
            bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (ok) {
                m.timer.capturing = true;
                try {
                    body();
                } catch (const Error&) {
                    ok = false;
                }
                m.timer.capturing = false;
                const hipError_t e = hipStreamEndCapture(s, &g);
                ok = ok && e == hipSuccess && g &&
                     hipGraphInstantiate(&cg.g[kind], g, nullptr, nullptr, 0) == hipSuccess;
                if (g) (void)hipGraphDestroy(g);
            }
            if (!ok || m.graph_epoch != before) {  // a capture must not reallocate
                m.drop_cell_graphs();
                m.graphs_off = !ok;
                (void)hipGetLastError();
                body();
                sync(s);
                return;
            }
        }
        if (replay && cg.g[kind]) HIPCHK(hipGraphLaunch(cg.g[kind], s));
        else body();
        sync(s);
    };
    double v = 0;
    const float* xa0 = xa;
    while (it < max_iter) {
        iteration(it == 0 ? 0 : xa == xa0 ? 1 : 2, xa, xb);
        std::memcpy(&r, m.h_words, 4);
        std::memcpy(&v, m.h_words + 2, sizeof(double));
        std::swap(xa, xb);
        it++;
        if (r < tol) break;
        if (watch.update(r) && m.cfg.stall_stop) break;
    }
    // growth during an eager call is adopted before any capture; after
    // one, it retires the graphs (they point into the old slots)
    if (cg.epoch != m.graph_epoch) {
        if (cg.g[0] || cg.g[1] || cg.g[2]) m.drop_cell_graphs();
        else cg.epoch = m.graph_epoch;
    }
    if (it == 0) {  // no iteration ran: the packed start, unpacked
        prologue();
        fill_u32(s, reinterpret_cast<unsigned*>(vb), 2, 0u);
        cell_violation(s, p.n_slices, p.slices, xa, p.eref_t, vb, m.cfg.g, m.cfg.norm_mean, m.cfg.norm_std,
                       m.cfg.norm_depth, m.cfg.energy_r_max);
        cell_unpack(s, p.M, xa, d_grid);
        d2h(m.h_words + 2, vb, sizeof(double), s);
        std::memcpy(&v, m.h_words + 2, sizeof(double));
    }
    if (iters) *iters = it;
    if (resid) *resid = r;
    if (violation) *violation = v;
}

}  // namespace
}  // namespace j2k

using namespace j2k;

extern "C" {

const char* j2k_last_error(void) { return g_err.c_str(); }

int j2k_device_count(int* n) {
    return guard([&] {
        need(n, "n");
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
        *n = c;
        if (c == 0) fail(J2K_ERR_NO_ADAPTER, "no HIP device is visible");
    });
}

int j2k_model_create(const uint8_t* st, size_t n, const char* cfg_json, int device, int precision, void** out) {
    return guard([&] {
        need(out, "out_model");
        *out = nullptr;
        *out = create_model(st, n, cfg_json, device, precision);
    });
}

int j2k_model_destroy(void* h) {
    return guard([&] { delete static_cast<Model*>(h); });
}

int j2k_model_config(const void* h, j2k_config* out) {
    return guard([&] {
        need(h, "model");
        need(out, "out");
        *out = static_cast<const Model*>(h)->cfg;
    });
}

int j2k_model_autotune(void* h, long long max_tokens, char* report, size_t report_len) {
    return guard([&] {
        Model& m = model_of(h);
        if (max_tokens < 0) fail(J2K_ERR_INVALID, "max_tokens must be >= 0");
        if (max_tokens > 0) autotune(m, (long)max_tokens);
        if (report && report_len) {
            const std::string j = tuning_json(m);
            const size_t n = std::min(j.size(), report_len - 1);
            std::memcpy(report, j.data(), n);
            report[n] = 0;
        }
    });
}

int j2k_model_timing(void* h, int enable, double* out) {
    return guard([&] {
        Model& m = model_of(h);
        DeviceGuard dg(m.device);
        m.timer.harvest();
        if (out)
            for (int c = 0; c < KC_COUNT; c++) {
                out[3 * c] = m.timer.ms[c];
                out[3 * c + 1] = m.timer.work[c];
                out[3 * c + 2] = m.timer.n[c];
            }
        if (enable >= 0) {
            m.timer.clear();
            m.timer.on = enable != 0;
        }
    });
}

int j2k_forward_batch_dev(void* h, int batch, const float* st, const float* meta, float* out, int add_state, void* s) {
    return guard([&] {
        Model& m = model_of(h);
        need_form(m, J2K_FORM_PATCH2D, "j2k_forward_batch_dev");
        need(st, "state");
        need(meta, "meta");
        need(out, "out");
        if (batch <= 0) fail(J2K_ERR_INVALID, "batch must be positive");
        DeviceGuard dg(m.device);
        forward_p(m, stream_of(m, s), batch, st, meta, out, add_state != 0);
    });
}

static int forward_host(void* h, const float* state, const float* meta, float* out, bool add) {
    return guard([&] {
        Model& m = model_of(h);
        need_form(m, J2K_FORM_PATCH2D, add ? "j2k_step" : "j2k_forward");
        need(state, "state");
        need(meta, "meta");
        need(out, "out");
        DeviceGuard dg(m.device);
        const j2k_config& c = m.cfg;
        const size_t cells = (size_t)c.channels * c.ny * c.nx;
        float* ds = m.get<float>(S_GRID, cells);
        float* dm = m.get<float>(S_META, c.meta_dim);
        float* dout = m.get<float>(S_WIN_A, cells);
        hipStream_t s = m.stream;
        h2d(ds, state, cells * 4, s);
        h2d(dm, meta, (size_t)c.meta_dim * 4, s);
        forward_p(m, s, 1, ds, dm, dout, add);
        d2h(out, dout, cells * 4, s);
    });
}

int j2k_forward(void* h, const float* state, const float* meta, float* incr) { return forward_host(h, state, meta, incr, false); }
int j2k_step(void* h, const float* state, const float* meta, float* next) { return forward_host(h, state, meta, next, true); }

int j2k_rollout_dev(void* h, float* x, const float* meta, int steps, int pm, int pe, double* etrace, void* s) {
    return guard([&] {
        Model& m = model_of(h);
        need_form(m, J2K_FORM_PATCH2D, "j2k_rollout_dev");
        need(x, "state");
        need(meta, "meta");
        if (steps < 0) fail(J2K_ERR_INVALID, "steps must be >= 0");
        DeviceGuard dg(m.device);
        rollout_dev(m, stream_of(m, s), x, meta, steps, pm, pe, etrace);
    });
}

int j2k_rollout(void* h, float* state_io, const float* meta, int steps, int pm, int pe, double* etrace, double* viol) {
    return guard([&] {
        Model& m = model_of(h);
        need_form(m, J2K_FORM_PATCH2D, "j2k_rollout");
        need(state_io, "state_io");
        need(meta, "meta");
        if (steps < 0) fail(J2K_ERR_INVALID, "steps must be >= 0");
        DeviceGuard dg(m.device);
        const j2k_config& c = m.cfg;
        const size_t cells = (size_t)c.channels * c.ny * c.nx;
        hipStream_t s = m.stream;
        float* dx = m.get<float>(S_GRID, cells);
        float* dm = m.get<float>(S_META, c.meta_dim);
        double* de = m.get<double>(S_TARG, (size_t)steps + 1);
        h2d(dx, state_io, cells * 4, s);
        h2d(dm, meta, (size_t)c.meta_dim * 4, s);
        rollout_dev(m, s, dx, dm, steps, pm, pe, de);
        std::vector<double> e((size_t)steps + 1);
        d2h(e.data(), de, e.size() * 8, s);
        d2h(state_io, dx, cells * 4, s);
        if (etrace) std::memcpy(etrace, e.data(), e.size() * 8);
        if (viol) *viol = violation_of(e, c.energy_r_max);
    });
}

int j2k_iterate_window(void* h, float* win, const float* meta, int T, int iters, int pm, int pe) {
    return guard([&] {
        Model& m = model_of(h);
        need_form(m, J2K_FORM_PATCH2D, "j2k_iterate_window");
        need(win, "window_io");
        need(meta, "meta");
        if (T < 1 || iters < 0) fail(J2K_ERR_INVALID, "need T >= 1 and iters >= 0");
        DeviceGuard dg(m.device);
        const j2k_config& c = m.cfg;
        const int n = c.ny * c.nx;
        const size_t g = 3 * (size_t)n;
        hipStream_t s = m.stream;
        float* wa = m.get<float>(S_WIN_A, g * T);
        float* wb = m.get<float>(S_WIN_B, g * T);
        float* dm = m.get<float>(S_META, c.meta_dim);
        double* sums = m.get<double>(S_SUMS, 2 * (size_t)T + 2);
        double* mass0 = sums;           // [T-1] copies of the slice-0 mass
        double* eprev = sums + T;       // [T-1] energies of slices 0..T-2
        h2d(wa, win, g * T * 4, s);
        h2d(dm, meta, (size_t)c.meta_dim * 4, s);
        HIPCHK(hipMemcpyAsync(wb, wa, g * 4, hipMemcpyDeviceToDevice, s));
        double m0 = 0;
        grid_sums(s, 1, n, wa, mass0, nullptr, c.g, c.norm_mean, c.norm_std, c.norm_depth);
        d2h(&m0, mass0, 8, s);
        std::vector<double> m0v((size_t)std::max(T - 1, 1), m0);
        h2d(mass0, m0v.data(), m0v.size() * 8, s);
        for (int it = 0; it < iters && T > 1; it++) {
            grid_sums(s, T - 1, n, wa, nullptr, eprev, c.g, c.norm_mean, c.norm_std, c.norm_depth);
            forward_p(m, s, T - 1, wa, dm, wb + g, true);
            project_p(s, T - 1, n, wb + g, mass0, pm, eprev, pe, nullptr, c.g, c.norm_mean, c.norm_std, c.norm_depth,
                      c.energy_r_max);
            std::swap(wa, wb);
        }
        d2h(win, wa, g * T * 4, s);
    });
}

static int grid_scalar(void* h, const float* x, int n, double* out, bool energy) {
    return guard([&] {
        Model& m = model_of(h);
        need(x, "x");
        need(out, "out");
        if (n <= 0) fail(J2K_ERR_INVALID, "n must be positive");
        DeviceGuard dg(m.device);
        const j2k_config& c = m.cfg;
        hipStream_t s = m.stream;
        float* dx = m.get<float>(S_GRID, 3 * (size_t)n);
        double* r = m.get<double>(S_SUMS, 2);
        h2d(dx, x, 12 * (size_t)n, s);
        grid_sums(s, 1, n, dx, r, r + 1, c.g, c.norm_mean, c.norm_std, c.norm_depth);
        double v[2];
        d2h(v, r, 16, s);
        *out = energy ? v[1] : v[0];
    });
}

int j2k_mass(void* h, const float* x, int n, double* out) { return grid_scalar(h, x, n, out, false); }
int j2k_energy(void* h, const float* x, int n, double* out) { return grid_scalar(h, x, n, out, true); }

static int project_host(void* h, float* x, int n, double value, bool energy) {
    return guard([&] {
        Model& m = model_of(h);
        need(x, "x");
        if (n <= 0) fail(J2K_ERR_INVALID, "n must be positive");
        DeviceGuard dg(m.device);
        const j2k_config& c = m.cfg;
        hipStream_t s = m.stream;
        float* dx = m.get<float>(S_GRID, 3 * (size_t)n);
        double* v = m.get<double>(S_SUMS, 2);
        h2d(dx, x, 12 * (size_t)n, s);
        h2d(v, &value, 8, s);
        project_p(s, 1, n, dx, v, energy ? 0 : 1, v, energy ? 1 : 0, nullptr, c.g, c.norm_mean, c.norm_std,
                  c.norm_depth, c.energy_r_max);
        d2h(x, dx, 12 * (size_t)n, s);
    });
}

int j2k_project_mass(void* h, float* x, int n, float target) { return project_host(h, x, n, (double)target, false); }
int j2k_project_energy(void* h, float* x, int n, float prev) { return project_host(h, x, n, (double)prev, true); }

int j2k_fill_blank(const float* grid, const uint8_t* known, int T, int ny, int nx, float* out) {
    return guard([&] {
        need(grid, "grid");
        need(known, "known");
        need(out, "out");
        if (T <= 0 || ny <= 0 || nx <= 0) fail(J2K_ERR_SHAPE, "grid dimensions must be positive");
        const size_t n = (size_t)ny * nx, M = (size_t)T * n;
        std::memcpy(out, grid, 3 * M * sizeof(float));
        std::vector<int> full;
        for (int s = 0; s < T; s++) {
            bool all = true;
            for (size_t i = 0; i < n && all; i++) all = known[s * n + i] != 0;
            if (all) full.push_back(s);
        }
        double gsum[3] = {0, 0, 0};
        long gcnt = 0;
        for (size_t q = 0; q < M; q++)
            if (known[q]) {
                for (int c = 0; c < 3; c++) gsum[c] += grid[c * M + q];
                gcnt++;
            }
        for (int s = 0; s < T; s++) {
            int src = -1, best = 1 << 30;
            for (int f : full)
                if (std::abs(f - s) < best) {
                    best = std::abs(f - s);
                    src = f;
                }
            float fillv[3] = {0, 0, 0};
            if (src < 0) {
                double ss[3] = {0, 0, 0};
                long cnt = 0;
                for (size_t i = 0; i < n; i++)
                    if (known[s * n + i]) {
                        for (int c = 0; c < 3; c++) ss[c] += grid[c * M + s * n + i];
                        cnt++;
                    }
                for (int c = 0; c < 3; c++)
                    fillv[c] = cnt ? (float)(ss[c] / cnt) : (gcnt ? (float)(gsum[c] / gcnt) : 0.0f);
            }
            for (size_t i = 0; i < n; i++) {
                const size_t q = s * n + i;
                if (known[q]) continue;
                for (int c = 0; c < 3; c++) out[c * M + q] = src >= 0 ? grid[c * M + (size_t)src * n + i] : fillv[c];
            }
        }
    });
}

int j2k_iterate_dev(void* h, float* d_grid, const uint8_t* d_known, const float* d_meta, int T, int ny, int nx,
                    int max_iter, float tol, int* iters, float* resid, double* violation, void* s) {
    return guard([&] {
        Model& m = model_of(h);
        need(d_grid, "grid");
        need(d_known, "known");
        need(d_meta, "meta");
        DeviceGuard dg(m.device);
        iterate_dev(m, stream_of(m, s), d_grid, d_known, d_meta, T, ny, nx, max_iter, tol, iters, resid, violation);
    });
}

int j2k_iterate(void* h, float* grid_io, const uint8_t* known, const float* meta, int T, int ny, int nx, int max_iter,
                float tol, int* iters, float* resid, double* violation) {
    return guard([&] {
        Model& m = model_of(h);
        need(grid_io, "grid_io");
        need(known, "known");
        need(meta, "meta");
        need_form(m, J2K_FORM_CELL_ST, "j2k_iterate");
        check_cell_dims(m, CellDims{T, ny, nx});
        DeviceGuard dg(m.device);
        const size_t M = (size_t)T * ny * nx;
        hipStream_t s = m.stream;
        float* dg_ = m.get<float>(S_GRID, 3 * M);
        uint8_t* dk = m.get<uint8_t>(S_KNOWN, M);
        float* dm = m.get<float>(S_META, m.cfg.meta_dim);
        h2d(dg_, grid_io, 12 * M, s);
        h2d(dk, known, M, s);
        h2d(dm, meta, (size_t)m.cfg.meta_dim * 4, s);
        iterate_dev(m, s, dg_, dk, dm, T, ny, nx, max_iter, tol, iters, resid, violation);
        d2h(grid_io, dg_, 12 * M, s);
    });
}

int j2k_cell_forward_dev(void* h, const float* d_grid, const uint8_t* d_known, const float* d_meta, int T, int ny,
                         int nx, float* d_y, void* st) {
    return guard([&] {
        Model& m = model_of(h);
        need(d_grid, "grid");
        need(d_known, "known");
        need(d_meta, "meta");
        need(d_y, "y");
        DeviceGuard dg(m.device);
        hipStream_t s = stream_of(m, st);
        CellRun run;
        cell_setup(m, s, CellDims{T, ny, nx}, d_meta, run);
        float* x4 = m.get<float>(S_X0, 4 * (size_t)run.pack.M);
        cell_pack(s, run.pack.M, d_grid, d_known, x4);
        cell_model(m, s, run.pack, x4, d_y);
    });
}

int j2k_score_dev(void* h, int batch, const float* d_x, const float* d_h, const float* d_u, const float* d_v,
                  const float* centers, int max_lag, j2k_eval_score* out, void* st) {
    return guard([&] {
        Model& m = model_of(h);
        need_form(m, J2K_FORM_PATCH2D, "j2k_score_dev");
        if (batch < 0) fail(J2K_ERR_INVALID, "batch must be >= 0");
        if (max_lag < 0) fail(J2K_ERR_INVALID, "max_lag must be >= 0");
        if (batch == 0) return;
        need(d_x, "x");
        need(d_h, "h");
        need(d_u, "u");
        need(d_v, "v");
        need(centers, "centers");
        need(out, "out");
        DeviceGuard dg(m.device);
        score_batch(m, stream_of(m, st), batch, d_x, d_h, d_u, d_v, centers, max_lag, out);
    });
}

int j2k_score(void* h, int batch, const float* x, const float* hh, const float* u, const float* v,
              const float* centers, int max_lag, j2k_eval_score* out) {
    return guard([&] {
        Model& m = model_of(h);
        need_form(m, J2K_FORM_PATCH2D, "j2k_score");
        if (batch < 0) fail(J2K_ERR_INVALID, "batch must be >= 0");
        if (max_lag < 0) fail(J2K_ERR_INVALID, "max_lag must be >= 0");
        if (batch == 0) return;
        need(x, "x");
        need(hh, "h");
        need(u, "u");
        need(v, "v");
        need(centers, "centers");
        need(out, "out");
        DeviceGuard dg(m.device);
        hipStream_t s = m.stream;
        const size_t n = (size_t)m.cfg.nx * m.cfg.ny, cells = (size_t)batch * n;
        float* dx = m.get<float>(S_X0, 3 * cells);
        float* ds = m.get<float>(S_XA, 3 * cells);
        h2d(dx, x, 3 * cells * 4, s);
        h2d(ds, hh, cells * 4, s);
        h2d(ds + cells, u, cells * 4, s);
        h2d(ds + 2 * cells, v, cells * 4, s);
        score_batch(m, s, batch, dx, ds, ds + cells, ds + 2 * cells, centers, max_lag, out);
    });
}

int j2k_rollout_truth_dev(void* h, void* solver, float* d_x, const float* d_meta, float* d_h, float* d_u, float* d_v,
                          int steps, float* d_corr, void* st) {
    return guard([&] {
        Model& m = model_of(h);
        need_form(m, J2K_FORM_PATCH2D, "j2k_rollout_truth_dev");
        need(solver, "solver");
        need(d_x, "x");
        need(d_meta, "meta");
        need(d_h, "h");
        need(d_u, "u");
        need(d_v, "v");
        if (steps < 0) fail(J2K_ERR_INVALID, "steps must be >= 0");
        if (steps > 0) need(d_corr, "corr");
        DeviceGuard dg(m.device);
        rollout_truth_dev(m, stream_of(m, st), solver, d_x, d_meta, d_h, d_u, d_v, steps, d_corr);
    });
}

int j2k_rollout_truth(void* h, void* solver, float* x_io, const float* meta, float* hh, float* u, float* v, int steps,
                      float* corr) {
    return guard([&] {
        Model& m = model_of(h);
        need_form(m, J2K_FORM_PATCH2D, "j2k_rollout_truth");
        need(solver, "solver");
        need(x_io, "x");
        need(meta, "meta");
        need(hh, "h");
        need(u, "u");
        need(v, "v");
        if (steps < 0) fail(J2K_ERR_INVALID, "steps must be >= 0");
        if (steps > 0) need(corr, "corr");
        DeviceGuard dg(m.device);
        const j2k_config& c = m.cfg;
        const size_t n = (size_t)c.nx * c.ny;
        hipStream_t s = m.stream;
        float* dx = m.get<float>(S_GRID, 3 * n);
        float* dm = m.get<float>(S_META, c.meta_dim);
        float* ds = m.get<float>(S_XA, 3 * n + (size_t)steps + 1);
        float* dc = ds + 3 * n;
        h2d(dx, x_io, 3 * n * 4, s);
        h2d(dm, meta, (size_t)c.meta_dim * 4, s);
        h2d(ds, hh, n * 4, s);
        h2d(ds + n, u, n * 4, s);
        h2d(ds + 2 * n, v, n * 4, s);
        rollout_truth_dev(m, s, solver, dx, dm, ds, ds + n, ds + 2 * n, steps, dc);
        if (steps > 0) d2h(corr, dc, (size_t)steps * 4, s);
        d2h(x_io, dx, 3 * n * 4, s);
        d2h(hh, ds, n * 4, s);
        d2h(u, ds + n, n * 4, s);
        d2h(v, ds + 2 * n, n * 4, s);
    });
}

}  // extern "C"
