// The device-side model: weights in HBM, one stream, growable scratch, and
// the forward passes both forms are built from.
#pragma once

#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "loader.h"

namespace j2k {

struct LayerW {
    float *n1g, *n1b, *wqkv, *bqkv, *wp, *bp, *n2g, *n2b, *w1, *b1, *w2, *b2;
    uint16_t *wqkv_t, *wp_t, *w1_t, *w2_t;  // bf16 copies, transposed to [N, K]
    // the same in MFMA fragment order (the row-panel kernel's operands; d = 128)
    uint16_t *wqkv_f = nullptr, *wp_f = nullptr, *w1_f = nullptr, *w2_f = nullptr;
    // FP8 form: E4M3 codes [N, K], each weight tensor's scale per output
    // column (the inputs' scales are measured per token and launch).
    uint8_t *wqkv_8, *wp_8, *w1_8, *w2_8;
    float *cs_qkv, *cs_p, *cs_1, *cs_2;
};

// Scratch slots, each grown on demand (the device is drained before a regrow).
enum Slot : int {
    S_X, S_H, S_HB, S_QKV, S_O, S_F, S_XN, S_T0, S_T1, S_Z0, S_Z2, S_C1, S_PATCH, S_MVEC,
    S_WIN_A, S_WIN_B, S_SUMS, S_SEGS, S_TILES, S_SLICES, S_X0, S_XA, S_XB, S_TARG, S_RESID, S_GRID, S_KNOWN, S_META,
    S_TOKREQ, S_CTILES, S_SPLITK, S_AMAX, S_COUNT
};

// Kernel classes timed with HIP events when timing is on (the bench's
// roofline figures): work is FLOPs for GEMM, algorithmic bytes otherwise.
// The row-panel kernel counts its activations read and written once:
// KC_ROWS when it also makes the next block's QKV, KC_ROWS_EMB when it
// does so for the first block from the embedding, KC_ROWS_TAIL otherwise.
enum KClass : int {
    KC_GEMM = 0,
    KC_ATTN = 1,
    KC_NORM = 2,
    KC_PROJ = 3,
    KC_ROWS = 4,
    KC_ROWS_TAIL = 5,
    KC_ROWS_EMB = 6,
    KC_COUNT = 7
};

struct KTimer {
    bool on = false;
    bool capturing = false;  // no event records inside a captured graph
    struct Rec {
        hipEvent_t a, b;
        int cls;
        double work;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double ms[KC_COUNT] = {}, work[KC_COUNT] = {}, n[KC_COUNT] = {};
    hipEvent_t take();
    void harvest();
    void clear();
    ~KTimer();
};

// Launch choices picked by the autotuner per bucket of packed tokens
// (bucket b holds M in [2^b, 2^(b+1))); -1 keeps the built-in default.
enum TuneOp : int {
    TO_GEMM_QKV = 0,   // block height of the qkv GEMM
    TO_GEMM_PROJ = 1,  // block height of the attention-output GEMM
    TO_GEMM_FF = 2,    // block height of ff1 / ff2 when the MLP is not fused
    TO_ATTN = 3,       // staging mode of the tiled cell attention (KvMode: joint, split, plane)
    TO_ROWS = 4,       // d = 128 token-local work: 0 apart (proj / qkv GEMMs), fused in one of the
                       // row-panel launch shapes 1 .. 4; d = 256 MLP: 0 two tile GEMMs, 1 one fused launch
    TO_COUNT = 5
};
struct Tuning {
    static constexpr int NB = 31;
    signed char pick[TO_COUNT][NB];
    static constexpr int NC = 5;  // candidates per op at most
    float us[TO_COUNT][NB][NC];   // measured µs of each candidate (0: not a candidate)
    int tuned_to = -1;          // highest bucket tuned so far
    bool enabled = true;        // iterate / server tune on first use (J2K_AUTOTUNE=0 turns it off)
    Tuning();
    static int bucket(long M);
    int choice(int op, long M) const;  // -1 when untuned
};

struct Model;
// Brackets the launches of its scope with events on `s` when timing is on.
struct TScope {
    Model& m;
    hipStream_t s;
    int cls;
    double work;
    hipEvent_t a = nullptr;
    TScope(Model& m, hipStream_t s, int cls, double work);
    ~TScope();
};

struct Model {
    j2k_config cfg{};
    KTimer timer;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> owned;
    float *embed_w = nullptr, *embed_b = nullptr, *pos = nullptr, *meta_w = nullptr, *meta_b = nullptr;
    float *norm_g = nullptr, *norm_b = nullptr;
    float *tm_w = nullptr, *tm_b = nullptr, *tc_w = nullptr, *tc_b = nullptr, *c1_w = nullptr, *c1_b = nullptr,
          *c2_w = nullptr, *c2_b = nullptr;
    float *ro_w = nullptr, *ro_b = nullptr;
    std::vector<LayerW> L;
    void* slot[S_COUNT] = {};
    size_t cap[S_COUNT] = {};
    int p_tables_batch = -1;  // batch size the patch2d seg/tile tables were built for
    // Grid the single-request cell_st tables in S_SEGS / S_TILES / S_SLICES /
    // S_CTILES were built for (T = -1: none; whoever else writes those slots
    // resets it), with their counts.
    struct {
        int T = -1, ny = 0, nx = 0, M = 0, n_slices = 0, n_tiles = 0, n_ctiles = 0;
    } cell_tables;
    int p_tiles = 0;
    Tuning tune;
    // Replayed single-request cell_st iterations (j2k_iterate_dev): the
    // first iteration with the request's set-up (meta vector, pack,
    // targets), then one graph per ping-pong parity for the later ones;
    // each ends with the unpack into the caller's grid and the copy of the
    // residual and violation words to h_words. Keyed by the grid's shape,
    // the caller's buffers and the scratch slots they were captured with.
    // graph_epoch moves whenever a scratch slot is reallocated or the tuning
    // changes, which retires them; the first call with a key runs eagerly so
    // that every slot has its size before a capture.
    struct {
        int T = -1, ny = 0, nx = 0;
        const void *x0 = nullptr, *xa = nullptr, *xb = nullptr;
        const void *grid = nullptr, *known = nullptr, *meta = nullptr;
        unsigned epoch = 0;
        int calls = 0;
        hipGraphExec_t g[3] = {nullptr, nullptr, nullptr};  // first, later from xa, later from xb
    } cell_graph;
    unsigned graph_epoch = 0;
    // pinned host words for the per-iteration residual and the violation
    // (device -> host copies into pageable memory take a staging detour)
    unsigned* h_words = nullptr;
    bool graphs_off = false;  // J2K_GRAPHS=0 at creation, or a capture failed: run eagerly
    void drop_cell_graphs();

    // A fixed region that scratch() carves every slot from while `on`,
    // instead of allocating (the allocation-free entry point): 256-byte
    // aligned, first come first served, J2K_ERR_OOM once it is exhausted.
    // The slots of the arena are separate from the model's own.
    struct {
        bool on = false;
        char* base = nullptr;
        size_t size = 0, used = 0;
        void* slot[S_COUNT] = {};
        size_t cap[S_COUNT] = {};
    } arena;

    ~Model();
    void* scratch(int s, size_t bytes);
    // The slot's current device buffer (the arena's while it is on).
    void* current(int s) const { return arena.on ? arena.slot[s] : slot[s]; }
    template <class T>
    T* get(int s, size_t count) {
        return static_cast<T*>(scratch(s, count * sizeof(T)));
    }
    AttnGeo geo() const;
    // Activations stored as bf16 (the bf16 form); bf16x2 keeps them f32 and
    // splits them into a bf16 pair inside the GEMM.
    bool bf16() const { return cfg.precision == J2K_PREC_BF16; }
    bool split() const { return cfg.precision == J2K_PREC_BF16X2; }
    bool fp8() const { return cfg.precision == J2K_PREC_FP8; }
};

// Makes `m.device` current for the scope of a call.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev);
    ~DeviceGuard();
};

Model* create_model(const uint8_t* st, size_t n, const char* cfg_json, int device, int precision);

// Looks up, once per device and with that device current, the device address
// of the allocation-free entry point's fixed scratch region (a failure is
// kept and reported by that entry point).
void noalloc_init(int device);

// The transformer trunk (cpu.rs:287-313) over M packed tokens: X [M, dim] is
// the residual stream. `ready` says what the caller has made of the first
// block's input: X only (LN1 is computed here), X and H = LN1(X), or X and
// the first block's QKV (the served bf16 forms). Leaves the stream
// un-normalised (callers apply the final norm), unless `ro` asks for the
// cell_st readout and the served d = 128 trunk takes it (then X's final
// value may never be stored): returns whether it did.
// With emb (ready == TRUNK_QKV, and only where rows_embed_in_trunk says
// so), X holds nothing yet: the first block's row-panel launch makes the
// embedding itself and stores X after its MLP.
enum { TRUNK_X = 0, TRUNK_H = 1, TRUNK_QKV = 2 };
bool trunk(Model& m, hipStream_t s, int M, float* X, int ready, const AttnPlan& ap, const Readout* ro = nullptr,
           const RowsEmbed* emb = nullptr);
// Whether the trunk over M tokens runs its first block in the row-panel
// launch that can take the embedding (served bf16 forms, d = 128, more than
// one block, the four-wave shape chosen for M).
bool rows_embed_in_trunk(const Model& m, int M);

// patch2d forward over B grids [B, 3, ny, nx] (cpu.rs:249-352); with
// add_state the result is state + forward (cpu.rs:84-88).
void forward_p(Model& m, hipStream_t s, int B, const float* d_state, const float* d_meta, float* d_out, bool add_state);

// Times the candidate launch shapes of the low-precision trunk on synthetic
// packs of 2^b tokens for every bucket b up to that of max_tokens (from
// 2^12; at most 2^20) not yet tuned, and keeps the fastest per bucket. Runs
// on the model's stream and synchronises; never call it inside a capture.
// Results do not depend on the choices (they only change the launch shape).
void autotune(Model& m, long max_tokens);
// Sets fixed choices for every bucket and turns tuning off ("gemm_bm=64",
// "attn_kv_mode=1", "rows_fused=1", comma separated; unnamed ops keep their
// defaults).
void force_tuning(Model& m, const char* spec);
// The tuning table as JSON text.
std::string tuning_json(const Model& m);

// Tile table for a list of segments (host side).
void build_tiles(const std::vector<Seg>& segs, int tpb, std::vector<Tile>& tiles);

}  // namespace j2k
