// Model construction (weights to HBM) and the two forward passes' shared parts.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "model.h"

#include <cstring>

namespace j2k {

DeviceGuard::DeviceGuard(int dev) {
    HIPCHK(hipGetDevice(&prev));
    if (prev != dev) HIPCHK(hipSetDevice(dev));
}
DeviceGuard::~DeviceGuard() {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
}

void Model::drop_cell_graphs() {
    for (hipGraphExec_t& g : cell_graph.g)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    cell_graph.T = -1;
    cell_graph.grid = cell_graph.known = cell_graph.meta = nullptr;
    cell_graph.calls = 0;
}

Model::~Model() {
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    drop_cell_graphs();
    if (h_words) (void)hipHostFree(h_words);
    for (void* p : owned) (void)hipFree(p);
    for (int i = 0; i < S_COUNT; i++)
        if (slot[i]) (void)hipFree(slot[i]);
    if (stream) (void)hipStreamDestroy(stream);
    if (prev >= 0) (void)hipSetDevice(prev);
}

void* Model::scratch(int s, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (arena.on) {  // carve from the caller's fixed region; never allocate
        if (arena.cap[s] >= bytes) return arena.slot[s];
        const size_t want = (bytes + 255) & ~(size_t)255;
        if (want > arena.size - arena.used)
            fail(J2K_ERR_OOM, "allocation-free scratch exhausted: " + std::to_string(arena.used) + " of " +
                                  std::to_string(arena.size) + " bytes in use, " + std::to_string(want) + " more asked");
        arena.slot[s] = arena.base + arena.used;
        arena.cap[s] = want;
        arena.used += want;
        return arena.slot[s];
    }
    if (cap[s] >= bytes) return slot[s];
    graph_epoch++;  // captured graphs may point into the old slot
    if (slot[s]) {
        // work on any stream (the model's, or a caller's passed to a _dev
        // entry point) may still read the old buffer
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipFree(slot[s]));
        slot[s] = nullptr;
        cap[s] = 0;
    }
    size_t want = bytes + bytes / 4;  // headroom so small growth does not reallocate every call
    hipError_t e = hipMalloc(&slot[s], want);
    if (e != hipSuccess) {
        slot[s] = nullptr;
        fail(J2K_ERR_OOM, "device scratch of " + std::to_string(want) + " bytes: " + hipGetErrorString(e));
    }
    cap[s] = want;
    return slot[s];
}

hipEvent_t KTimer::take() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    HIPCHK(hipEventCreate(&e));
    return e;
}

void KTimer::harvest() {
    for (Rec& r : recs) {
        HIPCHK(hipEventSynchronize(r.b));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
        ms[r.cls] += t;
        work[r.cls] += r.work;
        n[r.cls] += 1;
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    recs.clear();
}

void KTimer::clear() {
    harvest();
    for (int i = 0; i < KC_COUNT; i++) ms[i] = work[i] = n[i] = 0;
}

KTimer::~KTimer() {
    for (Rec& r : recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

TScope::TScope(Model& mm, hipStream_t ss, int c, double w) : m(mm), s(ss), cls(c), work(w) {
    if (m.timer.on && !m.timer.capturing) {
        a = m.timer.take();
        HIPCHK(hipEventRecord(a, s));
    }
}

TScope::~TScope() {
    if (!a) return;
    hipEvent_t b = m.timer.take();
    if (hipEventRecord(b, s) == hipSuccess) m.timer.recs.push_back(KTimer::Rec{a, b, cls, work});
}

AttnGeo Model::geo() const {
    AttnGeo g{};
    if (cfg.form == J2K_FORM_CELL_ST) {
        g.kind = ATTN_CELL_ST;
    } else {
        g.kind = cfg.mask_kind == J2K_MASK_FULL ? ATTN_FULL : ATTN_PATCH_STENCIL;
    }
    g.radius = cfg.radius;
    g.dilation = cfg.dilation;
    g.time_radius = cfg.time_radius;
    return g;
}

namespace {

uint16_t host_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

struct Uploader {
    Model& m;
    template <class T>
    T* put(const T* host, size_t count) {
        void* p = nullptr;
        hip_check(hipMalloc(&p, count * sizeof(T) + 16), "hipMalloc(weights)", J2K_ERR_OOM);
        m.owned.push_back(p);
        hip_check(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy(weights)", J2K_ERR_MAP);
        return static_cast<T*>(p);
    }
    float* tensor(const Weights& w, const std::string& name, std::vector<int64_t> shape) {
        const Tensor& t = get(w, name);
        if (t.shape != shape) {
            std::string want, have;
            for (auto s : shape) want += std::to_string(s) + ",";
            for (auto s : t.shape) have += std::to_string(s) + ",";
            fail(J2K_ERR_SHAPE, "tensor `" + name + "` has shape [" + have + "], expected [" + want + "]");
        }
        return put(t.data.data(), t.data.size());
    }
    // E4M3 copy of weights [K, N_i] joined along N, transposed to [N, K]:
    // each tensor keeps the codes and scale it was stored with (F8_E4M3
    // checkpoints) or gets scale amax / 448 and round-to-nearest-even codes.
    // cs receives the scale of every output column.
    uint8_t* f8_t(const std::vector<const Tensor*>& ts, int K, float** cs) {
        int N = 0;
        for (auto* t : ts) N += (int)t->shape[1];
        std::vector<uint8_t> codes((size_t)K * N);
        std::vector<float> col(N);
        int n0 = 0;
        for (auto* t : ts) {
            const int Ni = (int)t->shape[1];
            float sc = t->f8_scale;
            if (t->f8.empty()) {
                float amax = 0.f;
                for (float v : t->data) amax = std::max(amax, std::fabs(v));
                sc = amax > 0.f ? amax / 448.0f : 1.0f;
            }
            for (int k = 0; k < K; k++)
                for (int n = 0; n < Ni; n++) {
                    const size_t i = (size_t)k * Ni + n;
                    codes[(size_t)(n0 + n) * K + k] = t->f8.empty() ? e4m3_encode(t->data[i] / sc) : t->f8[i];
                }
            for (int n = 0; n < Ni; n++) col[n0 + n] = sc;
            n0 += Ni;
        }
        *cs = put(col.data(), col.size());
        return put(codes.data(), codes.size());
    }
    // [K, N] row-major f32 -> bf16 [N, K].
    uint16_t* bf16_t(const std::vector<float>& w, int K, int N) {
        std::vector<uint16_t> t((size_t)K * N);
        for (int k = 0; k < K; k++)
            for (int n = 0; n < N; n++) t[(size_t)n * K + k] = host_bf16(w[(size_t)k * N + n]);
        return put(t.data(), t.size());
    }
    // W [K, N] as bf16 in MFMA fragment order: block (nt, ks) of 16 rows n and
    // 32 columns k of W^T is 64 lanes x 8 values, lane l holding
    // W^T[16 nt + l%16][32 ks + 8 (l/16) .. +8]; blocks run ks-fastest. One
    // wave's fragment load is then 1 KB in a row.
    uint16_t* bf16_frag(const std::vector<float>& w, int K, int N) {
        if (N % 16 || K % 32) fail(J2K_ERR_LIMIT, "fragment order needs N % 16 == 0 and K % 32 == 0");
        std::vector<uint16_t> t((size_t)K * N);
        const int ks_n = K / 32;
        for (int nt = 0; nt < N / 16; nt++)
            for (int ks = 0; ks < ks_n; ks++)
                for (int l = 0; l < 64; l++)
                    for (int e = 0; e < 8; e++) {
                        const int n = 16 * nt + (l & 15), k = 32 * ks + 8 * (l >> 4) + e;
                        t[(((size_t)nt * ks_n + ks) * 64 + l) * 8 + e] = host_bf16(w[(size_t)k * N + n]);
                    }
        return put(t.data(), t.size());
    }
};

}  // namespace

Model* create_model(const uint8_t* st, size_t n, const char* cfg_json, int device, int precision) {
    if (precision != J2K_PREC_F32 && precision != J2K_PREC_BF16 && precision != J2K_PREC_BF16X2 &&
        precision != J2K_PREC_FP8)
        fail(J2K_ERR_INVALID, "precision must be F32, BF16, BF16X2 or FP8");
    if (!cfg_json) fail(J2K_ERR_INVALID, "config_json is NULL");
    Weights w = read_safetensors(st, n);
    j2k_config cfg = parse_config(cfg_json, w, precision);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) fail(J2K_ERR_NO_ADAPTER, "no HIP device is visible");
    if (device < 0 || device >= count)
        fail(J2K_ERR_NO_DEVICE, "device " + std::to_string(device) + " out of range (" + std::to_string(count) + " visible)");
    DeviceGuard dg(device);
    auto m = std::make_unique<Model>();
    m->cfg = cfg;
    if (const char* e = std::getenv("J2K_AUTOTUNE")) m->tune.enabled = std::strcmp(e, "0") != 0;
    if (const char* e = std::getenv("J2K_TUNE_FORCE")) force_tuning(*m, e);
    if (const char* e = std::getenv("J2K_GRAPHS")) m->graphs_off = std::strcmp(e, "0") == 0;
    m->device = device;
    hip_check(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "hipStreamCreate", J2K_ERR_NO_DEVICE);
    hip_check(hipHostMalloc((void**)&m->h_words, 4 * sizeof(unsigned)), "hipHostMalloc", J2K_ERR_OOM);
    if (cfg.form == J2K_FORM_CELL_ST) {
        attn_cell_prepare();
        noalloc_init(device);  // the allocation-free entry point's region on this device
    }
    Uploader up{*m};
    const int d = cfg.dim, ff = cfg.ff_inner;
    if (cfg.dim / cfg.heads != 8 && cfg.dim / cfg.heads != 16 && cfg.dim / cfg.heads != 32 && cfg.dim / cfg.heads != 64)
        fail(J2K_ERR_LIMIT, "head size dim/heads must be 8, 16, 32 or 64");
    if (d > 1024) fail(J2K_ERR_LIMIT, "dim above 1024 is not supported");
    if (cfg.form == J2K_FORM_PATCH2D) {
        m->embed_w = up.tensor(w, "embed.weight", {cfg.patch_dim, d});
        m->embed_b = up.tensor(w, "embed.bias", {d});
        m->pos = up.tensor(w, "pos", {1, cfg.tokens, d});
    } else {
        m->embed_w = up.tensor(w, "embed.weight", {cfg.channels + 1, d});
        m->embed_b = up.tensor(w, "embed.bias", {d});
    }
    m->meta_w = up.tensor(w, "meta.weight", {cfg.meta_dim, d});
    m->meta_b = up.tensor(w, "meta.bias", {d});
    for (int i = 0; i < cfg.layers; i++) {
        const std::string p = "blocks." + std::to_string(i) + ".";
        LayerW L{};
        L.n1g = up.tensor(w, p + "norm1.gamma", {d});
        L.n1b = up.tensor(w, p + "norm1.beta", {d});
        L.n2g = up.tensor(w, p + "norm2.gamma", {d});
        L.n2b = up.tensor(w, p + "norm2.beta", {d});
        // Fused QKV weight [d, 3d]: row i = q.w[i] | k.w[i] | v.w[i] (gpu.rs:288-301 layout).
        const Tensor& q = get(w, p + "q.weight");
        const Tensor& k = get(w, p + "k.weight");
        const Tensor& v = get(w, p + "v.weight");
        for (const Tensor* t : {&q, &k, &v})
            if (t->shape != std::vector<int64_t>{d, d}) fail(J2K_ERR_SHAPE, "tensor `" + p + "{q,k,v}.weight` must be [dim, dim]");
        std::vector<float> wqkv((size_t)d * 3 * d), bqkv(3 * (size_t)d);
        for (int r = 0; r < d; r++) {
            std::memcpy(&wqkv[(size_t)r * 3 * d], &q.data[(size_t)r * d], d * 4);
            std::memcpy(&wqkv[(size_t)r * 3 * d + d], &k.data[(size_t)r * d], d * 4);
            std::memcpy(&wqkv[(size_t)r * 3 * d + 2 * d], &v.data[(size_t)r * d], d * 4);
        }
        const char* bn[3] = {"q.bias", "k.bias", "v.bias"};
        for (int j = 0; j < 3; j++) {
            const Tensor& b = get(w, p + bn[j]);
            if (b.shape != std::vector<int64_t>{d}) fail(J2K_ERR_SHAPE, "tensor `" + p + bn[j] + "` must be [dim]");
            std::memcpy(&bqkv[(size_t)j * d], b.data.data(), d * 4);
        }
        L.wqkv = up.put(wqkv.data(), wqkv.size());
        L.bqkv = up.put(bqkv.data(), bqkv.size());
        L.wp = up.tensor(w, p + "proj.weight", {d, d});
        L.bp = up.tensor(w, p + "proj.bias", {d});
        L.w1 = up.tensor(w, p + "ff1.weight", {d, ff});
        L.b1 = up.tensor(w, p + "ff1.bias", {ff});
        L.w2 = up.tensor(w, p + "ff2.weight", {ff, d});
        L.b2 = up.tensor(w, p + "ff2.bias", {d});
        if (precision == J2K_PREC_FP8) {
            const Tensor &pw = get(w, p + "proj.weight"), &w1 = get(w, p + "ff1.weight"), &w2 = get(w, p + "ff2.weight");
            L.wqkv_8 = up.f8_t({&q, &k, &v}, d, &L.cs_qkv);
            L.wp_8 = up.f8_t({&pw}, d, &L.cs_p);
            L.w1_8 = up.f8_t({&w1}, d, &L.cs_1);
            L.w2_8 = up.f8_t({&w2}, ff, &L.cs_2);
        } else if (precision != J2K_PREC_F32) {
            L.wqkv_t = up.bf16_t(wqkv, d, 3 * d);
            L.wp_t = up.bf16_t(get(w, p + "proj.weight").data, d, d);
            L.w1_t = up.bf16_t(get(w, p + "ff1.weight").data, d, ff);
            L.w2_t = up.bf16_t(get(w, p + "ff2.weight").data, ff, d);
            if (rows_fused_ok(d, ff)) {
                L.wqkv_f = up.bf16_frag(wqkv, d, 3 * d);
                L.wp_f = up.bf16_frag(get(w, p + "proj.weight").data, d, d);
                L.w1_f = up.bf16_frag(get(w, p + "ff1.weight").data, d, ff);
                L.w2_f = up.bf16_frag(get(w, p + "ff2.weight").data, ff, d);
            } else if (mlp_fused_ok(d, ff)) {
                L.wqkv_f = up.bf16_frag(wqkv, d, 3 * d);
                L.wp_f = up.bf16_frag(get(w, p + "proj.weight").data, d, d);
                L.w1_f = up.bf16_frag(get(w, p + "ff1.weight").data, d, ff);
                L.w2_f = up.bf16_frag(get(w, p + "ff2.weight").data, ff, d);
            }
        }
        m->L.push_back(L);
    }
    m->norm_g = up.tensor(w, "norm.gamma", {d});
    m->norm_b = up.tensor(w, "norm.beta", {d});
    if (cfg.form == J2K_FORM_PATCH2D) {
        const int hd = cfg.head_dim, c = cfg.channels;
        m->tm_w = up.tensor(w, "head.token_mix.weight", {d, d, 3, 3});
        m->tm_b = up.tensor(w, "head.token_mix.bias", {d});
        m->tc_w = up.tensor(w, "head.to_cells.weight", {hd, d, 1, 1});
        m->tc_b = up.tensor(w, "head.to_cells.bias", {hd});
        m->c1_w = up.tensor(w, "head.conv1.weight", {hd, hd + c, 3, 3});
        m->c1_b = up.tensor(w, "head.conv1.bias", {hd});
        m->c2_w = up.tensor(w, "head.conv2.weight", {c, hd, 3, 3});
        m->c2_b = up.tensor(w, "head.conv2.bias", {c});
    } else {
        m->ro_w = up.tensor(w, "readout.weight", {d, cfg.channels});
        m->ro_b = up.tensor(w, "readout.bias", {cfg.channels});
    }
    return m.release();
}

void build_tiles(const std::vector<Seg>& segs, int tpb, std::vector<Tile>& tiles) {
    tiles.clear();
    for (size_t i = 0; i < segs.size(); i++) {
        const int n = segs[i].T * segs[i].ny * segs[i].nx;
        for (int st = 0; st < n; st += tpb) tiles.push_back(Tile{(int)i, st});
    }
}

namespace {

// Attention over one packed batch: the LDS-tiled kernel for the cell_st
// stencil when its neighbourhood fits in LDS, the generic gather otherwise.
template <class Act>
void attend(Model& m, hipStream_t s, int M, const Act* QKV, Act* O, const AttnPlan& ap) {
    const int d = m.cfg.dim, H = m.cfg.heads;
    const AttnGeo geo = m.geo();
    // untuned (the f32 and FP8 forms, or tuning off): K then V, the fastest
    // mode at the 64 x 64 x 16 and 128 x 128 x 16 workloads
    int mode = m.tune.choice(TO_ATTN, M);
    if (mode < 0) mode = KV_SPLIT;
    if (!attn_cell_ok(d / H, geo.radius, geo.time_radius, sizeof(Act), mode)) mode = KV_JOINT;
    const bool tiled = ap.ctiles && geo.kind == ATTN_CELL_ST &&
                       attn_cell_ok(d / H, geo.radius, geo.time_radius, sizeof(Act), mode);
    if (tiled) {
        if (sizeof(Act) == 4)
            attention_cell_f32(s, (const float*)QKV, (float*)O, d, H, geo.radius, geo.time_radius, ap.segs, ap.ctiles,
                               ap.n_ctiles, mode);
        else
            attention_cell_bf16(s, (const uint16_t*)QKV, (uint16_t*)O, d, H, geo.radius, geo.time_radius, ap.segs,
                                ap.ctiles, ap.n_ctiles, mode);
        return;
    }
    const int tpb = attn_tokens_per_block(d);
    if (sizeof(Act) == 4)
        attention_f32(s, (const float*)QKV, (float*)O, d, H, geo, ap.segs, ap.tiles, ap.n_tiles, tpb);
    else
        attention_bf16(s, (const uint16_t*)QKV, (uint16_t*)O, d, H, geo, ap.segs, ap.tiles, ap.n_tiles, tpb);
}

// The trunk in the served forms: bf16 weights on the tiled MFMA GEMM, with
// activations in bf16 (Act = uint16_t) or kept f32 and split into bf16 pairs
// inside the GEMM (Act = float). With dim == 128 the token-local part of a
// block runs in the row-panel kernel, in one of two arrangements of the same
// arithmetic (the autotuner picks): fused, where one launch takes O to the
// next block's QKV (proj, LN2, MLP, LN1', qkv), or apart, where the proj and
// qkv GEMMs run on their own and the row kernel does the MLP and LN1'.
// Other widths use the tile GEMMs with LayerNorm epilogues or launches.
template <class Act>
bool trunk_lowp(Model& m, hipStream_t s, int M, float* X, int ready, const AttnPlan& ap, const Readout* ro,
                const RowsEmbed* emb) {
    constexpr bool split = sizeof(Act) == 4;
    const int d = m.cfg.dim, ff = m.cfg.ff_inner, nl = m.cfg.layers;
    Act* Hh = m.get<Act>(split ? S_H : S_HB, (size_t)M * d);
    Act* QKV = m.get<Act>(S_QKV, (size_t)M * 3 * d);
    Act* O = m.get<Act>(S_O, (size_t)M * d);
    const bool rows = rows_fused_ok(d, ff);
    Act* F = rows ? nullptr : m.get<Act>(S_F, (size_t)M * ff);  // the row kernel keeps it on chip
    const double g = 2.0 * M * d, ln_b = (double)M * d * (4 + sizeof(Act)), at_b = 4.0 * sizeof(Act) * M * d;
    const int rows_pick = rows ? m.tune.choice(TO_ROWS, M) : 0;
    // d = 256: everything between two attentions in one launch (the tuner's
    // pick 2), the MLP in one launch (1), or tile GEMMs and LayerNorm launches
    const int wpick = mlp_fused_ok(d, ff) && !m.L.empty() && m.L[0].w1_f ? m.tune.choice(TO_ROWS, M) : 0;
    const bool mlp = wpick == 1, full = wpick == 2;
    const bool fused = rows_pick >= ROWS_4W;
    const int rows_shape = fused ? rows_pick : ROWS_4W;
    if (emb && !(ready == TRUNK_QKV && fused && rows_shape == ROWS_4W && nl > 1))
        fail(J2K_ERR_INVALID, "trunk: the embedding enters only a four-wave fused first block");
    auto gemm = [&](int op, int N, int K, const Act* A, const uint16_t* Bt, const float* bias, Act* C, int epi,
                    const float* lg, const float* lb) {
        TScope t(m, s, KC_GEMM, g / d * N * K);
        const int bm = m.tune.choice(op, M);
        if (split)
            gemm_bf16x2_tile(s, M, N, K, (const float*)A, Bt, bias, (float*)C, X, epi, lg, lb, (float*)Hh, bm);
        else
            gemm_bf16_tile(s, M, N, K, (const uint16_t*)A, Bt, bias, (uint16_t*)C, X, epi, lg, lb, (uint16_t*)Hh, bm);
    };
    auto norm = [&](const float* lg, const float* lb) {
        TScope t(m, s, KC_NORM, ln_b);
        if (split) layernorm(s, M, d, X, lg, lb, (float*)Hh, nullptr);
        else layernorm(s, M, d, X, lg, lb, nullptr, (uint16_t*)Hh);
    };
    if (ready == TRUNK_X) norm(m.L[0].n1g, m.L[0].n1b);
    if (ready == TRUNK_QKV) {
    } else if (fused) {
        TScope t(m, s, KC_GEMM, g * 3 * d);
        RowsArgs a;
        a.Hin = Hh;
        a.Wqkv = m.L[0].wqkv_f;
        a.bqkv = m.L[0].bqkv;
        a.QKV = QKV;
        rows_qkv(s, M, d, split, a);
    } else {
        gemm(TO_GEMM_QKV, 3 * d, d, Hh, m.L[0].wqkv_t, m.L[0].bqkv, QKV, EPI_STORE, nullptr, nullptr);
    }
    for (int l = 0; l < nl; l++) {
        const LayerW& L = m.L[l];
        const LayerW* next = l + 1 < nl ? &m.L[l + 1] : nullptr;
        if (l > 0 && !fused && !full) gemm(TO_GEMM_QKV, 3 * d, d, Hh, L.wqkv_t, L.bqkv, QKV, EPI_STORE, nullptr, nullptr);
        {
            TScope t(m, s, KC_ATTN, at_b);
            attend<Act>(m, s, M, QKV, O, ap);
        }
        if (rows) {
            RowsArgs a;
            a.W1 = L.w1_f;
            a.b1 = L.b1;
            a.W2 = L.w2_f;
            a.b2 = L.b2;
            a.X = X;
            if (next) {
                a.n1g = next->n1g;
                a.n1b = next->n1b;
            }
            // algorithmic bytes: the panel's input, X read and written, the output
            double bytes = (double)M * d * (sizeof(Act) + 8.0);
            if (fused) {
                a.O = O;
                a.Wp = L.wp_f;
                a.bp = L.bp;
                a.n2g = L.n2g;
                a.n2b = L.n2b;
                if (next) {
                    a.Wqkv = next->wqkv_f;
                    a.bqkv = next->bqkv;
                    a.QKV = QKV;
                    bytes += (double)M * 3 * d * sizeof(Act);
                }
            } else {
                gemm(TO_GEMM_PROJ, d, d, O, L.wp_t, L.bp, nullptr, EPI_RESID_LN, L.n2g, L.n2b);
                a.Hin = Hh;
                a.Hout = Hh;
                if (next) bytes += (double)M * d * sizeof(Act);
            }
            if (!next && ro && fused) a.ro = *ro;
            int cls = a.QKV ? KC_ROWS : KC_ROWS_TAIL;
            if (l == 0 && emb) {  // 16 bytes of features per token instead of X
                a.emb = *emb;
                bytes -= (double)M * d * 4 - 16.0 * M;
                cls = KC_ROWS_EMB;
            }
            TScope t(m, s, cls, bytes);
            rows_fused(s, M, d, ff, split, a, rows_shape);
            continue;
        }
        if (full) {
            RowsArgs a;
            a.O = O;
            a.Wp = L.wp_f;
            a.bp = L.bp;
            a.n2g = L.n2g;
            a.n2b = L.n2b;
            a.W1 = L.w1_f;
            a.b1 = L.b1;
            a.W2 = L.w2_f;
            a.b2 = L.b2;
            a.X = X;
            double bytes = (double)M * d * (sizeof(Act) + 8.0);  // O in, X read and written
            if (next) {
                a.n1g = next->n1g;
                a.n1b = next->n1b;
                a.Wqkv = next->wqkv_f;
                a.bqkv = next->bqkv;
                a.QKV = QKV;
                bytes += 3.0 * M * d * sizeof(Act);
            }
            TScope t(m, s, next ? KC_ROWS : KC_ROWS_TAIL, bytes);
            rows256(s, M, d, ff, split, a);
            continue;
        }
        if (d == 128) {
            gemm(TO_GEMM_PROJ, d, d, O, L.wp_t, L.bp, nullptr, EPI_RESID_LN, L.n2g, L.n2b);
        } else {
            gemm(TO_GEMM_PROJ, d, d, O, L.wp_t, L.bp, nullptr, EPI_RESID, nullptr, nullptr);
            norm(L.n2g, L.n2b);
        }
        if (mlp) {
            RowsArgs a;
            a.Hin = Hh;
            a.W1 = L.w1_f;
            a.b1 = L.b1;
            a.W2 = L.w2_f;
            a.b2 = L.b2;
            a.X = X;
            {
                TScope t(m, s, KC_ROWS_TAIL, (double)M * d * (sizeof(Act) + 8.0));
                mlp_fused(s, M, d, ff, split, a);
            }
            if (next) norm(next->n1g, next->n1b);
            continue;
        }
        gemm(TO_GEMM_FF, ff, d, Hh, L.w1_t, L.b1, F, EPI_GELU, nullptr, nullptr);
        if (!next) {
            gemm(TO_GEMM_FF, d, ff, F, L.w2_t, L.b2, nullptr, EPI_RESID, nullptr, nullptr);
        } else if (d == 128) {
            gemm(TO_GEMM_FF, d, ff, F, L.w2_t, L.b2, nullptr, EPI_RESID_LN, next->n1g, next->n1b);
        } else {
            gemm(TO_GEMM_FF, d, ff, F, L.w2_t, L.b2, nullptr, EPI_RESID, nullptr, nullptr);
            norm(next->n1g, next->n1b);
        }
    }
    if (!ro || !rows) return false;
    // the fused readout's arithmetic whichever launch shape ran the blocks
    if (!fused) rows_readout(s, M, d, X, *ro);
    return true;
}

// The trunk in the FP8 form: activations f32 in HBM, the four block
// products on E4M3 codes (tile GEMM quantizing each row of its A operand
// with the scale max|row| / 448, measured by a reduction launch just
// before), attention, LayerNorm and residual in f32. d = 128 takes the
// LayerNorms in the GEMM epilogues.
void trunk_f8(Model& m, hipStream_t s, int M, float* X, int ready, const AttnPlan& ap) {
    const int d = m.cfg.dim, ff = m.cfg.ff_inner, nl = m.cfg.layers;
    float* Hh = m.get<float>(S_H, (size_t)M * d);
    float* QKV = m.get<float>(S_QKV, (size_t)M * 3 * d);
    float* O = m.get<float>(S_O, (size_t)M * d);
    float* F = m.get<float>(S_F, (size_t)M * ff);
    float* rmax = m.get<float>(S_AMAX, (size_t)M);  // the rows' maxima of the next product's input
    const double g = 2.0 * M * d, ln_b = (double)M * d * 8, at_b = 16.0 * M * d;
    const bool fuse_ln = d == 128;
    auto gemm = [&](int op, int N, int K, const float* A, const uint8_t* Bt, const float* cs, const float* bias,
                    float* C, int epi, const float* lg, const float* lb) {
        TScope t(m, s, KC_GEMM, g / d * N * K);
        row_amax_f32(s, M, K, A, rmax);
        gemm_f8_tile(s, M, N, K, A, rmax, Bt, cs, bias, C, X, epi, lg, lb, Hh, m.tune.choice(op, M));
    };
    auto norm = [&](const float* lg, const float* lb) {
        TScope t(m, s, KC_NORM, ln_b);
        layernorm(s, M, d, X, lg, lb, Hh, nullptr);
    };
    if (ready == TRUNK_X) norm(m.L[0].n1g, m.L[0].n1b);
    for (int l = 0; l < nl; l++) {
        const LayerW& L = m.L[l];
        const LayerW* next = l + 1 < nl ? &m.L[l + 1] : nullptr;
        gemm(TO_GEMM_QKV, 3 * d, d, Hh, L.wqkv_8, L.cs_qkv, L.bqkv, QKV, EPI_STORE, nullptr, nullptr);
        {
            TScope t(m, s, KC_ATTN, at_b);
            attend<float>(m, s, M, QKV, O, ap);
        }
        if (fuse_ln) {
            gemm(TO_GEMM_PROJ, d, d, O, L.wp_8, L.cs_p, L.bp, nullptr, EPI_RESID_LN, L.n2g, L.n2b);
        } else {
            gemm(TO_GEMM_PROJ, d, d, O, L.wp_8, L.cs_p, L.bp, nullptr, EPI_RESID, nullptr, nullptr);
            norm(L.n2g, L.n2b);
        }
        gemm(TO_GEMM_FF, ff, d, Hh, L.w1_8, L.cs_1, L.b1, F, EPI_GELU, nullptr, nullptr);
        if (!next) {
            gemm(TO_GEMM_FF, d, ff, F, L.w2_8, L.cs_2, L.b2, nullptr, EPI_RESID, nullptr, nullptr);
        } else if (fuse_ln) {
            gemm(TO_GEMM_FF, d, ff, F, L.w2_8, L.cs_2, L.b2, nullptr, EPI_RESID_LN, next->n1g, next->n1b);
        } else {
            gemm(TO_GEMM_FF, d, ff, F, L.w2_8, L.cs_2, L.b2, nullptr, EPI_RESID, nullptr, nullptr);
            norm(next->n1g, next->n1b);
        }
    }
}

}  // namespace

bool rows_embed_in_trunk(const Model& m, int M) {
    return (m.bf16() || m.split()) && m.cfg.dim == 128 && rows_fused_ok(m.cfg.dim, m.cfg.ff_inner) &&
           m.cfg.layers > 1 && !m.L.empty() && m.L[0].wqkv_f && m.tune.choice(TO_ROWS, M) == ROWS_4W;
}

bool trunk(Model& m, hipStream_t s, int M, float* X, int ready, const AttnPlan& ap, const Readout* ro,
           const RowsEmbed* emb) {
    if (emb && !(m.bf16() || m.split())) fail(J2K_ERR_INVALID, "trunk: the embedding needs a served bf16 form");
    if (ready == TRUNK_QKV && (m.cfg.precision == J2K_PREC_F32 || m.fp8()))
        fail(J2K_ERR_INVALID, "trunk: a ready QKV needs a served bf16 form");
    const int d = m.cfg.dim, ff = m.cfg.ff_inner;
    if (m.cfg.precision == J2K_PREC_F32) {
        float* Hh = m.get<float>(S_H, (size_t)M * d);
        float* QKV = m.get<float>(S_QKV, (size_t)M * 3 * d);
        float* O = m.get<float>(S_O, (size_t)M * d);
        float* F = m.get<float>(S_F, (size_t)M * ff);
        // split-K workspace for the products of few tiles (small packs)
        const size_t wsn = std::max({gemm_f32_workspace(M, 3 * d, d), gemm_f32_workspace(M, d, d),
                                     gemm_f32_workspace(M, ff, d), gemm_f32_workspace(M, d, ff)});
        float* ws = wsn ? m.get<float>(S_SPLITK, wsn) : nullptr;
        const double ln_b = (double)M * d * 8, at_b = 16.0 * M * d, g = 2.0 * M * d;
        for (int l = 0; l < m.cfg.layers; l++) {
            const LayerW& L = m.L[l];
            if (l > 0 || ready == TRUNK_X) {
                TScope t(m, s, KC_NORM, ln_b);
                layernorm(s, M, d, X, L.n1g, L.n1b, Hh, nullptr);
            }
            {
                TScope t(m, s, KC_GEMM, g * 3 * d);
                gemm_f32(s, M, 3 * d, d, Hh, d, L.wqkv, L.bqkv, QKV, 3 * d, EPI_STORE, ws, wsn);
            }
            {
                TScope t(m, s, KC_ATTN, at_b);
                attend<float>(m, s, M, QKV, O, ap);
            }
            {
                TScope t(m, s, KC_GEMM, g * d);
                gemm_f32(s, M, d, d, O, d, L.wp, L.bp, X, d, EPI_RESID, ws, wsn);
            }
            {
                TScope t(m, s, KC_NORM, ln_b);
                layernorm(s, M, d, X, L.n2g, L.n2b, Hh, nullptr);
            }
            {
                TScope t(m, s, KC_GEMM, g * ff);
                gemm_f32(s, M, ff, d, Hh, d, L.w1, L.b1, F, ff, EPI_GELU, ws, wsn);
            }
            {
                TScope t(m, s, KC_GEMM, g * ff);
                gemm_f32(s, M, d, ff, F, ff, L.w2, L.b2, X, d, EPI_RESID, ws, wsn);
            }
        }
        return false;
    }
    if (m.fp8()) trunk_f8(m, s, M, X, ready, ap);
    else if (m.split()) return trunk_lowp<float>(m, s, M, X, ready, ap, ro, emb);
    else return trunk_lowp<uint16_t>(m, s, M, X, ready, ap, ro, emb);
    return false;
}

void forward_p(Model& m, hipStream_t s, int B, const float* d_state, const float* d_meta, float* d_out, bool add_state) {
    const j2k_config& c = m.cfg;
    const int d = c.dim, tokens = c.tokens, M = B * tokens, hd = c.head_dim, cells = c.ny * c.nx;
    float* patches = m.get<float>(S_PATCH, (size_t)M * c.patch_dim);
    float* mvec = m.get<float>(S_MVEC, d);
    float* X = m.get<float>(S_X, (size_t)M * d);
    patchify(s, B, c.channels, c.ny, c.nx, c.patch, d_state, patches);
    meta_linear(s, c.meta_dim, d, d_meta, m.meta_w, m.meta_b, mvec);
    // split-K workspace of the embedding and the CellHead convolutions (few tiles at 256 tokens)
    const int B_cells = B * cells, B_tok = B * tokens;
    const size_t wsn = std::max({gemm_f32_workspace(M, d, c.patch_dim), gemm_f32_workspace(d, B_tok, d * 9),
                                 gemm_f32_workspace(hd, B_tok, d), gemm_f32_workspace(hd, B_cells, (hd + c.channels) * 9),
                                 gemm_f32_workspace(c.channels, B_cells, hd * 9)});
    float* ws = wsn ? m.get<float>(S_SPLITK, wsn) : nullptr;
    gemm_embed_p(s, M, d, c.patch_dim, patches, m.embed_w, m.embed_b, m.pos, tokens, mvec, X, ws, wsn);

    if (m.p_tables_batch != B) {
        std::vector<Seg> segs(B);
        for (int b = 0; b < B; b++) segs[b] = Seg{b * tokens, 1, c.tokens_y, c.tokens_x};
        std::vector<Tile> tiles;
        build_tiles(segs, attn_tokens_per_block(c.dim), tiles);
        HIPCHK(hipStreamSynchronize(s));
        Seg* ds = m.get<Seg>(S_SEGS, segs.size());
        Tile* dt = m.get<Tile>(S_TILES, tiles.size());
        HIPCHK(hipMemcpy(ds, segs.data(), segs.size() * sizeof(Seg), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dt, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice));
        m.p_tables_batch = B;
        m.cell_tables.T = -1;
        m.p_tiles = (int)tiles.size();
    }
    AttnPlan ap;
    ap.segs = static_cast<const Seg*>(m.current(S_SEGS));
    ap.tiles = static_cast<const Tile*>(m.current(S_TILES));
    ap.n_tiles = m.p_tiles;
    trunk(m, s, M, X, TRUNK_X, ap);

    // Final norm and CellHead (cpu.rs:315-349).
    float* xn = m.get<float>(S_XN, (size_t)M * d);
    float* t0 = m.get<float>(S_T0, (size_t)M * d);
    float* t1 = m.get<float>(S_T1, (size_t)M * d);
    float* z0 = m.get<float>(S_Z0, (size_t)B * hd * tokens);
    float* z2 = m.get<float>(S_Z2, (size_t)B * hd * cells);
    float* c1 = m.get<float>(S_C1, (size_t)B * hd * cells);
    layernorm(s, M, d, X, m.norm_g, m.norm_b, xn, nullptr);
    transpose_td(s, B, tokens, d, xn, t0);
    ws = wsn ? m.get<float>(S_SPLITK, wsn) : nullptr;  // the trunk may have grown the slot
    conv_f32(s, B, d, c.tokens_y, c.tokens_x, t0, nullptr, d, m.tm_w, m.tm_b, d, 3, t1, 1, t0, ws, wsn);
    conv_f32(s, B, d, c.tokens_y, c.tokens_x, t1, nullptr, d, m.tc_w, m.tc_b, hd, 1, z0, 0, nullptr, ws, wsn);
    upsample2(s, B, hd, c.tokens_y, c.tokens_x, c.patch, z0, z2, (long)hd * cells);
    conv_f32(s, B, hd + c.channels, c.ny, c.nx, z2, d_state, hd, m.c1_w, m.c1_b, hd, 3, c1, 1, nullptr, ws, wsn);
    conv_f32(s, B, hd, c.ny, c.nx, c1, nullptr, hd, m.c2_w, m.c2_b, c.channels, 3, d_out, 0,
             add_state ? d_state : nullptr, ws, wsn);
}

}  // namespace j2k
