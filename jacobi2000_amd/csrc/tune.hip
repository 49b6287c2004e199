// Autotuner for the launch shapes of the low-precision trunk.
//
// For each bucket of packed tokens (2^b .. 2^(b+1)) it builds a synthetic
// pack of 2^b tokens (one segment of T slices of 64 x 64 cells), times every
// candidate of every tunable launch on the model's own weights and stream
// with HIP events, and keeps the fastest. Candidates only change how work is
// split into blocks, never the arithmetic of an output element, so the
// choice cannot change a result:
//   qkv / proj / ff GEMMs   block height 64 or 128 rows
//   cell attention          K and V staged together, K then V, or a time
//                           plane at a time
//   d = 128 token-local     proj / qkv GEMMs apart, or fused with the MLP
//                           (4 waves of 32 columns or 8 of 16)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "model.h"

namespace j2k {

namespace {

constexpr int GEMM_BM[2] = {64, 128};
constexpr int B_MIN = 12, B_MAX = 20;  // 4,096 .. 1,048,576 tokens
constexpr int REPS = 3;

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    Events() {
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
    }
    ~Events() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Best-of-REPS µs of `launch` on stream s (one untimed run first).
template <class F>
float time_us(hipStream_t s, Events& ev, F&& launch) {
    launch();
    float best = INFINITY;
    for (int r = 0; r < REPS; r++) {
        HIPCHK(hipEventRecord(ev.a, s));
        launch();
        HIPCHK(hipEventRecord(ev.b, s));
        HIPCHK(hipEventSynchronize(ev.b));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev.a, ev.b));
        best = std::fmin(best, ms * 1000.f);
    }
    return best;
}

template <class Act>
void tune_bucket(Model& m, int b, Events& ev) {
    constexpr bool split = sizeof(Act) == 4;
    hipStream_t s = m.stream;
    const int d = m.cfg.dim, ff = m.cfg.ff_inner, H = m.cfg.heads;
    const long M = 1L << b;
    const LayerW& L = m.L[0];
    float* X = m.get<float>(S_X, (size_t)M * d);
    Act* Hh = m.get<Act>(split ? S_H : S_HB, (size_t)M * d);
    Act* QKV = m.get<Act>(S_QKV, (size_t)M * 3 * d);
    Act* O = m.get<Act>(S_O, (size_t)M * d);
    Act* F = rows_fused_ok(d, ff) ? nullptr : m.get<Act>(S_F, (size_t)M * ff);
    HIPCHK(hipMemsetAsync(X, 0, (size_t)M * d * sizeof(float), s));
    HIPCHK(hipMemsetAsync(Hh, 0, (size_t)M * d * sizeof(Act), s));
    HIPCHK(hipMemsetAsync(QKV, 0, (size_t)M * 3 * d * sizeof(Act), s));
    HIPCHK(hipMemsetAsync(O, 0, (size_t)M * d * sizeof(Act), s));
    if (F) HIPCHK(hipMemsetAsync(F, 0, (size_t)M * ff * sizeof(Act), s));

    auto gemm = [&](int N, int K, const Act* A, const uint16_t* Bt, const float* bias, Act* C, int epi,
                    const float* lg, const float* lb, int bm) {
        if (split)
            gemm_bf16x2_tile(s, (int)M, N, K, (const float*)A, Bt, bias, (float*)C, X, epi, lg, lb, (float*)Hh, bm);
        else
            gemm_bf16_tile(s, (int)M, N, K, (const uint16_t*)A, Bt, bias, (uint16_t*)C, X, epi, lg, lb,
                           (uint16_t*)Hh, bm);
    };
    const bool fuse_ln = d == 128;
    for (int c = 0; c < 2; c++) {
        const int bm = GEMM_BM[c];
        m.tune.us[TO_GEMM_QKV][b][c] =
            time_us(s, ev, [&] { gemm(3 * d, d, Hh, L.wqkv_t, L.bqkv, QKV, EPI_STORE, nullptr, nullptr, bm); });
        m.tune.us[TO_GEMM_PROJ][b][c] = time_us(s, ev, [&] {
            gemm(d, d, O, L.wp_t, L.bp, nullptr, fuse_ln ? EPI_RESID_LN : EPI_RESID, L.n2g, L.n2b, bm);
        });
        if (!rows_fused_ok(d, ff))
            m.tune.us[TO_GEMM_FF][b][c] = time_us(s, ev, [&] {
                gemm(ff, d, Hh, L.w1_t, L.b1, F, EPI_GELU, nullptr, nullptr, bm);
                gemm(d, ff, F, L.w2_t, L.b2, nullptr, EPI_RESID, nullptr, nullptr, bm);
            });
    }

    const AttnGeo geo = m.geo();
    if (geo.kind == ATTN_CELL_ST) {
        // one segment of T slices of 64 x 64 cells (M = 4096 T)
        std::vector<Seg> segs{Seg{0, (int)(M / 4096), 64, 64}};
        std::vector<CellTile> tiles;
        build_cell_tiles(segs, tiles);
        Seg* dsegs = m.get<Seg>(S_SEGS, 1);
        CellTile* dtiles = m.get<CellTile>(S_CTILES, tiles.size());
        HIPCHK(hipMemcpyAsync(dsegs, segs.data(), sizeof(Seg), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dtiles, tiles.data(), tiles.size() * sizeof(CellTile), hipMemcpyHostToDevice, s));
        m.p_tables_batch = -1;  // the patch2d and cell_st tables share these slots
        m.cell_tables.T = -1;
        for (int mode = 0; mode < KV_MODES; mode++) {
            if (!attn_cell_ok(d / H, geo.radius, geo.time_radius, sizeof(Act), mode)) continue;
            m.tune.us[TO_ATTN][b][mode] = time_us(s, ev, [&] {
                if (split)
                    attention_cell_f32(s, (const float*)QKV, (float*)O, d, H, geo.radius, geo.time_radius, dsegs,
                                       dtiles, (int)tiles.size(), mode);
                else
                    attention_cell_bf16(s, (const uint16_t*)QKV, (uint16_t*)O, d, H, geo.radius, geo.time_radius,
                                        dsegs, dtiles, (int)tiles.size(), mode);
            });
        }
    }
    auto settle = [&](int op) {
        const float* u = m.tune.us[op][b];
        int best = -1;
        for (int c = 0; c < Tuning::NC; c++)
            if (u[c] > 0.f && (best < 0 || u[c] < u[best])) best = c;
        if (best >= 0) m.tune.pick[op][b] = (signed char)best;
    };
    for (int op = 0; op < TO_COUNT; op++) settle(op);

    if (rows_fused_ok(d, ff) && m.cfg.layers > 1) {
        // the token-local work between two attentions, apart (with the block
        // heights just chosen) or fused into one row-panel launch
        const LayerW& N1 = m.L[1];
        const int bq = GEMM_BM[m.tune.pick[TO_GEMM_QKV][b]], bp = GEMM_BM[m.tune.pick[TO_GEMM_PROJ][b]];
        RowsArgs r;
        r.W1 = L.w1_f;
        r.b1 = L.b1;
        r.W2 = L.w2_f;
        r.b2 = L.b2;
        r.X = X;
        r.n1g = N1.n1g;
        r.n1b = N1.n1b;
        m.tune.us[TO_ROWS][b][0] = time_us(s, ev, [&] {
            RowsArgs a = r;
            a.Hin = Hh;
            a.Hout = Hh;
            gemm(d, d, O, L.wp_t, L.bp, nullptr, EPI_RESID_LN, L.n2g, L.n2b, bp);
            rows_fused(s, (int)M, d, ff, split, a);
            gemm(3 * d, d, Hh, N1.wqkv_t, N1.bqkv, QKV, EPI_STORE, nullptr, nullptr, bq);
        });
        for (int shape : {ROWS_4W, ROWS_8W, ROWS_4W_F64, ROWS_2X2})
            m.tune.us[TO_ROWS][b][shape] = time_us(s, ev, [&] {
                RowsArgs a = r;
                a.O = O;
                a.Wp = L.wp_f;
                a.bp = L.bp;
                a.n2g = L.n2g;
                a.n2b = L.n2b;
                a.Wqkv = N1.wqkv_f;
                a.bqkv = N1.bqkv;
                a.QKV = QKV;
                rows_fused(s, (int)M, d, ff, split, a, shape);
            });
        settle(TO_ROWS);
    }
    if (mlp_fused_ok(d, ff) && L.w1_f && m.cfg.layers > 1) {
        // the token-local work between two attentions (proj + residual, LN2,
        // MLP, LN1 of the next block, its qkv): tile GEMMs and LayerNorm
        // launches with the block heights just chosen (0), the MLP in one
        // launch (1), or all of it in one launch (2)
        const LayerW& N1 = m.L[1];
        const int bq = GEMM_BM[m.tune.pick[TO_GEMM_QKV][b]], bp = GEMM_BM[m.tune.pick[TO_GEMM_PROJ][b]];
        const int bf = GEMM_BM[m.tune.pick[TO_GEMM_FF][b]];
        auto norm = [&](const float* g, const float* be) {
            if (split) layernorm(s, (int)M, d, X, g, be, (float*)Hh, nullptr);
            else layernorm(s, (int)M, d, X, g, be, nullptr, (uint16_t*)Hh);
        };
        RowsArgs r;
        r.W1 = L.w1_f;
        r.b1 = L.b1;
        r.W2 = L.w2_f;
        r.b2 = L.b2;
        r.X = X;
        m.tune.us[TO_ROWS][b][0] = time_us(s, ev, [&] {
            gemm(d, d, O, L.wp_t, L.bp, nullptr, EPI_RESID, nullptr, nullptr, bp);
            norm(L.n2g, L.n2b);
            gemm(ff, d, Hh, L.w1_t, L.b1, F, EPI_GELU, nullptr, nullptr, bf);
            gemm(d, ff, F, L.w2_t, L.b2, nullptr, EPI_RESID, nullptr, nullptr, bf);
            norm(N1.n1g, N1.n1b);
            gemm(3 * d, d, Hh, N1.wqkv_t, N1.bqkv, QKV, EPI_STORE, nullptr, nullptr, bq);
        });
        m.tune.us[TO_ROWS][b][1] = time_us(s, ev, [&] {
            gemm(d, d, O, L.wp_t, L.bp, nullptr, EPI_RESID, nullptr, nullptr, bp);
            norm(L.n2g, L.n2b);
            RowsArgs a = r;
            a.Hin = Hh;
            mlp_fused(s, (int)M, d, ff, split, a);
            norm(N1.n1g, N1.n1b);
            gemm(3 * d, d, Hh, N1.wqkv_t, N1.bqkv, QKV, EPI_STORE, nullptr, nullptr, bq);
        });
        m.tune.us[TO_ROWS][b][2] = time_us(s, ev, [&] {
            RowsArgs a = r;
            a.O = O;
            a.Wp = L.wp_f;
            a.bp = L.bp;
            a.n2g = L.n2g;
            a.n2b = L.n2b;
            a.n1g = N1.n1g;
            a.n1b = N1.n1b;
            a.Wqkv = N1.wqkv_f;
            a.bqkv = N1.bqkv;
            a.QKV = QKV;
            rows256(s, (int)M, d, ff, split, a);
        });
        settle(TO_ROWS);
    }
}

}  // namespace

Tuning::Tuning() {
    std::memset(pick, -1, sizeof(pick));
    std::memset(us, 0, sizeof(us));
}

int Tuning::bucket(long M) {
    int b = 0;
    while (b + 1 < NB && (1L << (b + 1)) <= M) b++;
    return b;
}

int Tuning::choice(int op, long M) const {
    if (tuned_to < 0) return -1;
    int b = bucket(M);
    if (b > tuned_to) b = tuned_to;
    if (b < B_MIN) b = B_MIN;
    const int c = pick[op][b];
    if (c < 0) return -1;
    return op == TO_ATTN || op == TO_ROWS ? c : GEMM_BM[c];
}

void force_tuning(Model& m, const char* spec) {
    // "gemm_bm=64|128", "attn_kv_mode=0..4", "rows_fused=0..4", comma separated; every bucket
    int bm = -1, mode = -1, rows = -1;
    std::string t(spec);
    size_t pos = 0;
    while (pos <= t.size()) {
        size_t end = t.find(',', pos);
        if (end == std::string::npos) end = t.size();
        const std::string kv = t.substr(pos, end - pos);
        const size_t eq = kv.find('=');
        if (eq != std::string::npos) {
            const std::string k = kv.substr(0, eq);
            const int v = std::atoi(kv.c_str() + eq + 1);
            if (k == "gemm_bm" && (v == 64 || v == 128)) bm = v == 64 ? 0 : 1;
            else if (k == "attn_kv_mode" && v >= 0 && v < KV_MODES) mode = v;
            else if (k == "rows_fused" && v >= 0 && v <= ROWS_2X2) rows = v;
            else fail(J2K_ERR_INVALID, "J2K_TUNE_FORCE: unknown setting '" + kv + "'");
        }
        pos = end + 1;
    }
    for (int b = 0; b < Tuning::NB; b++) {
        if (bm >= 0) m.tune.pick[TO_GEMM_QKV][b] = m.tune.pick[TO_GEMM_PROJ][b] = m.tune.pick[TO_GEMM_FF][b] = bm;
        if (mode >= 0) m.tune.pick[TO_ATTN][b] = mode;
        if (rows >= 0) m.tune.pick[TO_ROWS][b] = rows;
    }
    m.tune.tuned_to = B_MAX;
    m.tune.enabled = false;
    m.graph_epoch++;
}

void autotune(Model& m, long max_tokens) {
    // the plain form keeps its fixed shapes; the FP8 form its default block heights
    if (m.cfg.precision == J2K_PREC_F32 || m.fp8()) return;
    int top = Tuning::bucket(max_tokens);
    if (top > B_MAX) top = B_MAX;
    if (top < B_MIN) top = B_MIN;
    if (top <= m.tune.tuned_to) return;
    DeviceGuard g(m.device);
    // The tuner rewrites the model's scratch and tables on its own stream:
    // first let every stream of the device finish, a caller's stream that
    // may still be reading them included.
    HIPCHK(hipDeviceSynchronize());
    const bool timing = m.timer.on;
    m.timer.on = false;
    Events ev;
    for (int b = std::max(B_MIN, m.tune.tuned_to + 1); b <= top; b++) {
        if (m.split()) tune_bucket<float>(m, b, ev);
        else tune_bucket<uint16_t>(m, b, ev);
    }
    HIPCHK(hipStreamSynchronize(m.stream));
    m.timer.on = timing;
    m.tune.tuned_to = top;
    m.graph_epoch++;  // captured iterations used the old choices
}

std::string tuning_json(const Model& m) {
    static const char* names[TO_COUNT] = {"gemm_qkv_bm", "gemm_proj_bm", "gemm_ff_bm", "attn_kv_mode", "rows_fused"};
    std::string out = "{\"tuned_to_tokens\": ";
    out += std::to_string(m.tune.tuned_to < 0 ? 0L : 1L << m.tune.tuned_to);
    out += ", \"buckets\": [";
    bool first = true;
    for (int b = B_MIN; b <= m.tune.tuned_to; b++) {
        char buf[200];
        std::snprintf(buf, sizeof buf, "%s{\"tokens\": %ld", first ? "" : ", ", 1L << b);
        out += buf;
        first = false;
        for (int op = 0; op < TO_COUNT; op++) {
            const int c = m.tune.pick[op][b];
            if (c < 0) continue;
            const int v = op == TO_ATTN || op == TO_ROWS ? c : GEMM_BM[c];
            if (op == TO_ROWS || op == TO_ATTN)
                std::snprintf(buf, sizeof buf, ", \"%s\": %d, \"%s_us\": [%.2f, %.2f, %.2f, %.2f, %.2f]", names[op],
                              v, names[op], m.tune.us[op][b][0], m.tune.us[op][b][1], m.tune.us[op][b][2],
                              m.tune.us[op][b][3], m.tune.us[op][b][4]);
            else
                std::snprintf(buf, sizeof buf, ", \"%s\": %d, \"%s_us\": [%.2f, %.2f]", names[op], v, names[op],
                              m.tune.us[op][b][0], m.tune.us[op][b][1]);
            out += buf;
        }
        out += "}";
    }
    out += "]}";
    return out;
}

}  // namespace j2k
