// Continuous-batching request loop for the cell_st form on one GPU.
//
// A worker thread owns the model's stream. Each turn it (1) admits queued
// requests while the active set fits the HBM budget and max_active, (2)
// repacks the active requests' tokens into one contiguous layout when the set
// changed, (3) runs one fixed-point iteration over all of them (one ragged
// launch per kernel per layer), replayed from a captured hipGraph while the
// packing is unchanged, and (4) retires every request whose own residual fell
// under tol, or that reached max_iter, copying its grid back to the host.
//
// A request's result does not depend on the mix: every kernel computes a
// token from its own row and its own request's neighbours, and every
// reduction (mass, energy, residual) runs per slice in a fixed order.
//
// A rollout request (j2k_server_submit_rollout) chains windows of T slices
// along time: window w's last slice becomes window w+1's known slice 0, its
// other slices start as copies of it, and the request stays active until
// `ticks` slices are covered. The next window is built on the device from
// that slice of the packed state (nothing goes back up from the host); each
// finished window is copied out into the trajectory. Each window retires on
// its own residual like a plain request, and its energy violation,
// iterations and completion time are kept per window.
//
// Host traffic per turn: the new requests' grids and masks go up into their
// own regions of one staging buffer and the retiring requests' grids come
// down from another, with the launches between them back to back and one
// synchronisation per phase, not per request. While an iteration runs, the
// pending requests are copied into the other of two upload buffers and
// uploaded on a copy stream, and the previous turn's retired grids, already
// downloaded on that stream, are copied into their requests.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "abi.h"
#include "cell.h"

namespace j2k {

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t id = 0;
    CellDims dims{};
    std::vector<float> grid;     // [3, T, ny, nx], blank cells filled; the result on retirement
    std::vector<uint8_t> known;  // [T, ny, nx]
    std::vector<float> meta;
    j2k_result res{};
    int iters = 0;
    float resid = 0.f;
    size_t bytes = 0;
    int base = -1;  // first token in the current packing (-1: not yet on the device)
    int stage_buf = -1;    // >= 0: grid and mask already copied into that pinned staging buffer
    size_t stage_off = 0;  // at this offset
    int slice0 = 0;  // its first slice in the packing's slice table
    StallWatch watch;
    // rollout requests: ticks > 0
    int ticks = 0, window = 0, wi = 0;  // total slices, slices per window, windows done
    int prev_T = 0;  // > 0: the next window starts from slice prev_T - 1 of the packed state at `base`
    std::vector<float> traj;            // [3, ticks, ny, nx]
    std::vector<double> win_done, win_viol;
    std::vector<int> win_iters;
    bool converged_all = true;
};

// Slices covered by window w of a rollout (windows overlap by one slice).
int window_slices(int ticks, int window, int w) { return std::min(window, ticks - w * (window - 1)); }
int window_count(int ticks, int window) { return (ticks - 1 + window - 2) / (window - 1); }

// Window input from an anchor slice [3, ny*nx]: every slice a copy of it,
// slice 0 known (the blank fill of a forward request).
void set_window(Request& r, const float* anchor, int T) {
    const size_t n = (size_t)r.dims.ny * r.dims.nx;
    r.dims.T = T;
    r.grid.resize(3 * T * n);
    r.known.assign(T * n, 0);
    std::fill(r.known.begin(), r.known.begin() + n, (uint8_t)1);
    for (int c = 0; c < 3; c++)
        for (int t = 0; t < T; t++) std::memcpy(&r.grid[(c * T + t) * n], anchor + c * n, n * sizeof(float));
    r.iters = 0;
    r.resid = 0.f;
    r.watch = StallWatch{};
    r.base = -1;
}

// Device bytes one token of an active request holds: three float4 state
// vectors, its request index, and its rows of every activation buffer.
size_t bytes_per_token(const Model& m) {
    const size_t d = m.cfg.dim, e = m.bf16() ? 2 : 4;
    const bool rows = (m.bf16() || m.split()) && rows_fused_ok(m.cfg.dim, m.cfg.ff_inner);
    const size_t ff = rows ? 0 : m.cfg.ff_inner;
    return 3 * 16 + 4 + d * 4 + d * e + 3 * d * e + d * e + ff * e;
}

}  // namespace

struct Server {
    Model& m;
    size_t budget;
    int max_active;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::unique_ptr<Request>> pending;
    std::vector<std::unique_ptr<Request>> active;  // owned by the worker thread once admitted
    std::map<uint64_t, std::unique_ptr<Request>> done;
    size_t n_active = 0;  // for drain(): active.size() as last published under mu
    uint64_t next_id = 1;
    bool stop = false;
    std::string worker_error;
    int worker_code = J2K_OK;
    std::thread worker;
    std::atomic<uint64_t> launches{0}, replays{0};
    int peak_active = 0;
    size_t peak_bytes = 0;

    // Packed device state: x0 (anchors), xc (current iterate), xn (next).
    // alt0/altc receive a repack, then swap with x0/xc.
    CellTables tab;
    CellPack pack;
    bool packed = false;
    float *x0 = nullptr, *xc = nullptr, *xn = nullptr, *alt0 = nullptr, *altc = nullptr;
    int* d_tokreq = nullptr;
    size_t cap_tok = 0;
    Seg* d_segs = nullptr;
    Tile* d_tiles = nullptr;
    CellTile* d_ctiles = nullptr;
    size_t cap_ctiles = 0;
    Slice* d_slices = nullptr;
    float* d_mvec = nullptr;
    unsigned* d_resid = nullptr;
    double* d_targ = nullptr;
    size_t cap_tab = 0;
    std::vector<void*> owned;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // [0]: xc->xn with the current naming, [1]: swapped
    int parity = 0;
    int since_repack = 0;  // iterations run on the current packing
    bool graphs_enabled = true;
    unsigned* h_resid = nullptr;  // pinned
    size_t h_resid_cap = 0;
    // staging of the host copies: uploads of new requests (grid then mask,
    // each request in its own 256-byte aligned region), downloads of retiring
    // ones, and the retiring requests' slice table and violation words
    unsigned char* d_in = nullptr;
    unsigned char* d_out = nullptr;
    size_t cap_in = 0, cap_out = 0;
    Slice* d_rsl = nullptr;
    size_t cap_rsl = 0;
    unsigned long long* d_viol = nullptr;
    size_t cap_viol = 0;
    // pinned host staging of the same copies, so each turn's traffic is one
    // DMA per direction; reserved at creation for the largest pack the
    // budget admits (up to 2^24 tokens), grown if a turn needs more
    // Uploads go through two pinned buffers: a repack uploads from
    // h_stage[up] and flips `up`; while the iteration after it runs, the
    // worker copies the next pending requests into the other buffer
    // (prestage), so their host copy is off the critical path.
    unsigned char* h_stage[2] = {nullptr, nullptr};
    size_t cap_stage[2] = {0, 0};
    int up = 0;
    std::vector<Request*> staged;  // pre-staged into h_stage[up]
    size_t staged_end = 0;
    unsigned char* h_out = nullptr;
    size_t cap_hout = 0;
    // pinned host copies of a turn's small tables (tiles, segments, slices,
    // meta vectors, pack and move jobs; at retirement the slice table, its
    // reference-energy indices and the unpack jobs), uploaded without a
    // synchronisation; and the device job tables
    unsigned char* h_tab = nullptr;
    size_t cap_htab = 0;
    CellJob* d_jobs = nullptr;
    size_t cap_jobs = 0;
    int* d_ridx = nullptr;
    size_t cap_ridx = 0;
    // Retirement: its tables go through h_rtab, its download runs on s_copy
    // behind the unpack, and the host copies (into a request's grid, or a
    // rollout window's slices into its trajectory) wait for the next turn's
    // iteration to be launched (finish_retire), so the download and the
    // copies overlap the GPU's next iteration. Until then finished requests
    // sit in `deferred`, counted as active for drain(); a rollout with
    // windows to go stays active (its next window is made on the device).
    unsigned char* h_rtab = nullptr;
    size_t cap_hrtab = 0;
    hipStream_t s_copy = nullptr;
    hipEvent_t ev_unpacked = nullptr, ev_out = nullptr;
    // Pre-staged requests are also uploaded on s_copy while the iteration
    // runs: d_in [0, uploaded_end) holds h_stage[up] [0, uploaded_end), once
    // ev_uploaded has passed. d_in is free for it after the turn's pack
    // launches (ev_packed).
    hipEvent_t ev_packed = nullptr, ev_uploaded = nullptr;
    size_t uploaded_end = 0;
    struct Landing {
        Request* r;
        size_t off;          // its grid in h_out
        int T, t0;           // its slices; t0 >= 0: a rollout window's first tick
        int iters;           // the iterations it took
        bool converged, last;  // last: the rollout's final window
    };
    std::vector<Landing> landings;  // in retirement order (the violation words' order)
    std::vector<std::unique_ptr<Request>> deferred;
    size_t deferred_bytes = 0;
    // J2K_SERVER_PHASES=1 at creation: seconds the worker spent per phase,
    // printed to stderr when the server is destroyed
    bool phases_on = false;
    double ph_repack = 0, ph_iter = 0, ph_finish = 0, ph_turns = 0;
    std::string ph_detail;  // per repack: tables, staging copies, packing launches, scratch, final wait (ms)

    Server(Model& model, size_t b, int ma)
        : m(model), budget(b), max_active(ma), graphs_enabled(!model.graphs_off) {
        const char* e = std::getenv("J2K_SERVER_PHASES");
        phases_on = e && std::strcmp(e, "0") != 0;
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
        if (phases_on)
            std::fprintf(stderr, "j2k server phases: %.0f turns, repack %.4f s, iterations %.4f s, retire %.4f s%s\n",
                         ph_turns, ph_repack, ph_iter, ph_finish, ph_detail.c_str());
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(m.device);
        (void)hipStreamSynchronize(m.stream);
        if (s_copy) {
            (void)hipStreamSynchronize(s_copy);
            (void)hipStreamDestroy(s_copy);
        }
        for (hipEvent_t e : {ev_unpacked, ev_out, ev_packed, ev_uploaded})
            if (e) (void)hipEventDestroy(e);
        drop_graphs();
        for (void* p : owned) (void)hipFree(p);
        if (h_resid) (void)hipHostFree(h_resid);
        for (unsigned char* p : h_stage)
            if (p) (void)hipHostFree(p);
        if (h_out) (void)hipHostFree(h_out);
        if (h_tab) (void)hipHostFree(h_tab);
        if (h_rtab) (void)hipHostFree(h_rtab);
        if (prev >= 0) (void)hipSetDevice(prev);
    }

    template <class T>
    T* dev_alloc(size_t count) {
        void* p = nullptr;
        hip_check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), "server hipMalloc", J2K_ERR_OOM);
        owned.push_back(p);
        return static_cast<T*>(p);
    }
    void dev_free(void* p) {
        if (!p) return;
        auto it = std::find(owned.begin(), owned.end(), p);
        if (it != owned.end()) owned.erase(it);
        (void)hipFree(p);
    }
    static size_t aligned(size_t b) { return (b + 255) & ~(size_t)255; }
    // Grows a server-owned device buffer to at least `bytes` (contents not kept).
    template <class T>
    void reserve(T*& p, size_t& cap, size_t count) {
        if (count <= cap) return;
        HIPCHK(hipStreamSynchronize(m.stream));
        dev_free(p);
        cap = count + count / 4;
        p = dev_alloc<T>(cap);
    }

    // Grows a pinned host staging buffer to at least `bytes` (contents not kept).
    void reserve_host(unsigned char*& p, size_t& cap, size_t bytes) {
        if (bytes <= cap) return;
        HIPCHK(hipStreamSynchronize(m.stream));
        if (p) HIPCHK(hipHostFree(p));
        p = nullptr;
        cap = 0;
        const size_t want = bytes + bytes / 4;
        hip_check(hipHostMalloc((void**)&p, want), "server staging hipHostMalloc", J2K_ERR_OOM);
        cap = want;
    }

    // Host copies of one turn, (dst, src, bytes) each, spread over up to
    // eight threads in pieces of at most 4 MB when there is much to copy.
    struct Copy {
        void* dst;
        const void* src;
        size_t bytes;
    };
    static void copy_all(const std::vector<Copy>& jobs) {
        size_t total = 0;
        for (const Copy& c : jobs) total += c.bytes;
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const int nt = total >= (size_t(16) << 20) ? (int)std::min(8u, hw) : 1;
        if (nt <= 1) {
            for (const Copy& c : jobs) std::memcpy(c.dst, c.src, c.bytes);
            return;
        }
        constexpr size_t piece = size_t(4) << 20;
        std::vector<Copy> pieces;
        for (const Copy& c : jobs)
            for (size_t o = 0; o < c.bytes; o += piece)
                pieces.push_back(Copy{static_cast<char*>(c.dst) + o, static_cast<const char*>(c.src) + o,
                                      std::min(piece, c.bytes - o)});
        std::atomic<size_t> next{0};
        auto work = [&] {
            for (size_t i = next++; i < pieces.size(); i = next++) std::memcpy(pieces[i].dst, pieces[i].src, pieces[i].bytes);
        };
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
    }

    void drop_graphs() {
        for (auto& g : graph)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
    }

    // Rebuild the packing for `active`: surviving requests carry x0 and the
    // current iterate to their new place, new requests are uploaded and packed.
    void repack() {
        hipStream_t s = m.stream;
        HIPCHK(hipStreamSynchronize(s));
        drop_graphs();
        std::vector<CellDims> dims;
        for (auto& r : active) dims.push_back(r->dims);
        const bool multi = active.size() > 1;
        const double r0 = now_s();
        // tok_req is made on the device; the generic gather's tiles only when it may run
        tab.build(dims, attn_tokens_per_block(m.cfg.dim), false, cell_needs_generic_tiles(m));
        const double r1 = now_s();
        const size_t M = (size_t)tab.M;
        const bool grow = M > cap_tok;
        const size_t want = grow ? M + M / 4 : cap_tok;
        if (grow) {
            dev_free(alt0);
            dev_free(altc);
            alt0 = dev_alloc<float>(4 * want);
            altc = dev_alloc<float>(4 * want);
        }
        {  // every new grid and mask into pinned staging (the pre-staged ones are there already): one upload
            size_t end = staged_end;
            for (auto& r : active)
                if (r->base < 0 && r->stage_buf != up) end += aligned(13 * (size_t)r->dims.tokens());
            if (end > cap_stage[up]) {  // grows (contents lost): stage every new request again
                // an upload from the old buffer on s_copy ends before it is freed
                HIPCHK(hipEventSynchronize(ev_uploaded));
                for (Request* r : staged) r->stage_buf = -1;
                staged.clear();
                staged_end = 0;
                uploaded_end = 0;
                end = 0;
                for (auto& r : active)
                    if (r->base < 0) end += aligned(13 * (size_t)r->dims.tokens());
                reserve_host(h_stage[up], cap_stage[up], end);
            }
            unsigned char* hb = h_stage[up];
            std::vector<Copy> jobs;
            size_t o = staged_end;
            for (auto& r : active)
                if (r->base < 0 && r->stage_buf != up) {
                    const size_t T = (size_t)r->dims.tokens();
                    jobs.push_back(Copy{hb + o, r->grid.data(), 12 * T});
                    jobs.push_back(Copy{hb + o + 12 * T, r->known.data(), T});
                    r->stage_buf = up;
                    r->stage_off = o;
                    o += aligned(13 * T);
                }
            copy_all(jobs);
            // the uploads already made on s_copy end before d_in may be
            // replaced, and before the pack launches read it
            HIPCHK(hipEventSynchronize(ev_uploaded));
            if (end > cap_in) uploaded_end = 0;
            reserve(d_in, cap_in, end);
            if (end > uploaded_end)
                HIPCHK(hipMemcpyAsync(d_in + uploaded_end, hb + uploaded_end, end - uploaded_end, hipMemcpyHostToDevice, s));
            uploaded_end = 0;
            // pre-staged requests left pending are staged again for a later turn
            for (Request* r : staged) r->stage_buf = -1;
            staged.clear();
            staged_end = 0;
            up ^= 1;
        }
        const double r2 = now_s();
        // new requests are packed, survivors moved, by one launch each over
        // a job table; a rollout's next window is made from its last slice
        std::vector<CellJob> jobs;  // packs, then moves
        size_t n_pack = 0;
        int max_pack = 0, max_move = 0;
        {
            for (size_t i = 0; i < active.size(); i++) {
                Request& r = *active[i];
                const int T = r.dims.tokens();
                if (r.base < 0) {
                    jobs.push_back(CellJob{(long)r.stage_off, tab.segs[i].base, T});
                    r.stage_buf = -1;
                    max_pack = std::max(max_pack, T);
                }
            }
            n_pack = jobs.size();
            for (size_t i = 0; i < active.size(); i++) {
                const Request& r = *active[i];
                if (r.base >= 0 && r.prev_T == 0) {
                    jobs.push_back(CellJob{(long)r.base, tab.segs[i].base, r.dims.tokens()});
                    max_move = std::max(max_move, r.dims.tokens());
                }
            }
        }
        const size_t nseg = tab.segs.size(), ntile = tab.tiles.size(), nsl = tab.slices.size();
        const size_t md = (size_t)m.cfg.meta_dim, nreq = active.size();
        // the turn's small tables in one pinned buffer: ctiles, segs, tiles, slices, metas, jobs
        size_t off[7] = {0};
        {
            const size_t bytes[6] = {tab.ctiles.size() * sizeof(CellTile), nseg * sizeof(Seg), ntile * sizeof(Tile),
                                     nsl * sizeof(Slice), md * nreq * sizeof(float), jobs.size() * sizeof(CellJob)};
            for (int i = 0; i < 6; i++) off[i + 1] = off[i] + aligned(bytes[i]);
            reserve_host(h_tab, cap_htab, off[6]);
            std::memcpy(h_tab + off[0], tab.ctiles.data(), bytes[0]);
            std::memcpy(h_tab + off[1], tab.segs.data(), bytes[1]);
            std::memcpy(h_tab + off[2], tab.tiles.data(), bytes[2]);
            std::memcpy(h_tab + off[3], tab.slices.data(), bytes[3]);
            for (size_t i = 0; i < nreq; i++)
                std::memcpy(h_tab + off[4] + i * md * sizeof(float), active[i]->meta.data(), md * sizeof(float));
            std::memcpy(h_tab + off[5], jobs.data(), bytes[5]);
        }
        reserve(d_jobs, cap_jobs, jobs.size());
        if (!jobs.empty())
            HIPCHK(hipMemcpyAsync(d_jobs, h_tab + off[5], jobs.size() * sizeof(CellJob), hipMemcpyHostToDevice, s));
        cell_pack_jobs(s, (int)n_pack, d_jobs, max_pack, d_in, alt0, altc);
        HIPCHK(hipEventRecord(ev_packed, s));  // d_in is free for the next turn's uploads after this
        cell_move_jobs(s, (int)(jobs.size() - n_pack), d_jobs + n_pack, max_move, x0, xc, alt0, altc);
        int slices = 0;  // the slice table lists each request's slices in order
        for (size_t i = 0; i < active.size(); i++) {
            Request& r = *active[i];
            const size_t nb = (size_t)tab.segs[i].base;
            if (r.base >= 0 && r.prev_T > 0) {  // a rollout's next window, from its last slice on the device
                const size_t n = (size_t)r.dims.ny * r.dims.nx;
                cell_window(s, (int)n, r.dims.T, xc + 4 * ((size_t)r.base + (size_t)(r.prev_T - 1) * n), alt0 + 4 * nb,
                            altc + 4 * nb);
                r.prev_T = 0;
            }
            r.base = (int)nb;
            r.slice0 = slices;
            slices += r.dims.T;
        }
        std::swap(x0, alt0);
        std::swap(xc, altc);
        if (grow) {
            HIPCHK(hipStreamSynchronize(s));
            dev_free(alt0);
            dev_free(altc);
            dev_free(xn);
            dev_free(d_tokreq);
            alt0 = dev_alloc<float>(4 * want);
            altc = dev_alloc<float>(4 * want);
            xn = dev_alloc<float>(4 * want);
            d_tokreq = dev_alloc<int>(want);
            cap_tok = want;
        }
        parity = 0;
        since_repack = 0;

        const size_t need = std::max({nseg, ntile, nsl, (size_t)1});
        if (need > cap_tab) {
            for (void* p : {(void*)d_segs, (void*)d_tiles, (void*)d_slices, (void*)d_mvec, (void*)d_resid, (void*)d_targ})
                dev_free(p);
            const size_t w = need + need / 2;
            d_segs = dev_alloc<Seg>(w);
            d_tiles = dev_alloc<Tile>(w);
            d_slices = dev_alloc<Slice>(w);
            d_mvec = dev_alloc<float>(w * m.cfg.dim);
            d_resid = dev_alloc<unsigned>(w + 4);
            d_targ = dev_alloc<double>(2 * w);
            cap_tab = w;
        }
        if (tab.ctiles.size() > cap_ctiles) {
            dev_free(d_ctiles);
            cap_ctiles = tab.ctiles.size() + tab.ctiles.size() / 2;
            d_ctiles = dev_alloc<CellTile>(cap_ctiles);
        }
        HIPCHK(hipMemcpyAsync(d_ctiles, h_tab + off[0], tab.ctiles.size() * sizeof(CellTile), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_segs, h_tab + off[1], nseg * sizeof(Seg), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_tiles, h_tab + off[2], ntile * sizeof(Tile), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_slices, h_tab + off[3], nsl * sizeof(Slice), hipMemcpyHostToDevice, s));
        if (multi) seg_tokens(s, (int)nseg, d_segs, d_tokreq);
        float* dmeta = m.get<float>(S_META, md * nreq);
        HIPCHK(hipMemcpyAsync(dmeta, h_tab + off[4], md * nreq * sizeof(float), hipMemcpyHostToDevice, s));
        meta_linear_n(s, (int)nreq, (int)md, m.cfg.dim, dmeta, m.meta_w, m.meta_b, d_mvec);
        pack.M = (int)M;
        pack.n_req = (int)active.size();
        pack.n_slices = (int)nsl;
        pack.n_tiles = (int)ntile;
        pack.segs = d_segs;
        pack.tiles = d_tiles;
        pack.ctiles = d_ctiles;
        pack.n_ctiles = (int)tab.ctiles.size();
        pack.slices = d_slices;
        pack.tok_req = multi ? d_tokreq : nullptr;
        pack.mvec = d_mvec;
        pack.mass_t = d_targ;
        pack.eref_t = d_targ + nsl;
        pack.resid = d_resid;
        // targets depend on each request's x0 only, so recomputing them is exact
        cell_targets(s, (int)nsl, d_slices, x0, pack.mass_t, pack.eref_t, m.cfg.g, m.cfg.norm_mean, m.cfg.norm_std,
                     m.cfg.norm_depth);
        const double r3 = now_s();
        // grow the packed activation scratch now, never inside a capture
        const size_t d = m.cfg.dim, ff = m.cfg.ff_inner, e = m.bf16() ? 2 : 4;
        m.get<float>(S_X, M * d);
        if (m.bf16()) m.get<uint16_t>(S_HB, M * d);
        else m.get<float>(S_H, M * d);
        m.get<char>(S_QKV, M * 3 * d * e);
        m.get<char>(S_O, M * d * e);
        if (!((m.bf16() || m.split()) && rows_fused_ok(m.cfg.dim, m.cfg.ff_inner))) m.get<char>(S_F, M * ff * e);
        m.p_tables_batch = -1;
        m.cell_tables.T = -1;
        if (active.size() > h_resid_cap) {
            if (h_resid) HIPCHK(hipHostFree(h_resid));
            h_resid = nullptr;
            h_resid_cap = active.size() + 64;
            hip_check(hipHostMalloc((void**)&h_resid, h_resid_cap * sizeof(unsigned)), "hipHostMalloc", J2K_ERR_OOM);
        }
        const double r4 = now_s();
        HIPCHK(hipStreamSynchronize(s));
        packed = true;
        if (phases_on) {
            char b[160];
            std::snprintf(b, sizeof b, "; repack of %zu tokens: tables %.1f, staging %.1f, packing %.1f, scratch %.1f, wait %.1f", M,
                          1e3 * (r1 - r0), 1e3 * (r2 - r1), 1e3 * (r3 - r2), 1e3 * (r4 - r3), 1e3 * (now_s() - r4));
            ph_detail += b;
        }
    }

    // While the GPU runs an iteration: copy the next pending requests (in
    // admission order, up to max_active and the buffer's size) into the
    // staging buffer the next repack uploads from. Only the worker admits
    // requests, so the ones it takes here stay valid after the lock.
    bool prestage() {
        std::vector<Request*> take;
        {
            std::lock_guard<std::mutex> lk(mu);
            size_t end = staged_end;
            size_t n = staged.size();
            for (auto& p : pending) {
                if ((int)n >= max_active) break;
                Request* r = p.get();
                if (r->stage_buf == up) continue;
                const size_t b = aligned(13 * (size_t)r->dims.tokens());
                if (end + b > cap_stage[up]) break;
                r->stage_buf = up;
                r->stage_off = end;
                end += b;
                n++;
                take.push_back(r);
            }
            staged_end = end;
        }
        std::vector<Copy> jobs;
        for (Request* r : take) {
            const size_t T = (size_t)r->dims.tokens();
            jobs.push_back(Copy{h_stage[up] + r->stage_off, r->grid.data(), 12 * T});
            jobs.push_back(Copy{h_stage[up] + r->stage_off + 12 * T, r->known.data(), T});
        }
        copy_all(jobs);
        staged.insert(staged.end(), take.begin(), take.end());
        // and on to the device while the iteration runs, when d_in holds them
        if (!take.empty() && staged_end <= cap_in && uploaded_end < staged_end) {
            HIPCHK(hipStreamWaitEvent(s_copy, ev_packed, 0));
            HIPCHK(hipMemcpyAsync(d_in + uploaded_end, h_stage[up] + uploaded_end, staged_end - uploaded_end,
                                  hipMemcpyHostToDevice, s_copy));
            HIPCHK(hipEventRecord(ev_uploaded, s_copy));
            uploaded_end = staged_end;
        }
        return !take.empty();
    }

    // One iteration xc -> xn, then the names swap.
    void run_iteration() {
        hipStream_t s = m.stream;
        // The first iteration on a new packing runs eagerly: a packing that
        // changes every turn (requests converging in one iteration) would
        // otherwise pay a capture per turn that is never replayed. From the
        // second iteration on the same packing, each parity is captured once
        // and replayed.
        if (graphs_enabled && since_repack >= 1 && !graph[parity]) {
            hipGraph_t g = nullptr;
            bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (ok) {
                m.timer.capturing = true;
                try {
                    cell_iteration(m, s, pack, x0, xc, xn);
                } catch (const Error&) {
                    ok = false;
                }
                m.timer.capturing = false;
                hipError_t e = hipStreamEndCapture(s, &g);
                ok = ok && e == hipSuccess && g && hipGraphInstantiate(&graph[parity], g, nullptr, nullptr, 0) == hipSuccess;
                if (g) (void)hipGraphDestroy(g);
            }
            if (!ok) {
                graph[parity] = nullptr;
                graphs_enabled = false;
                (void)hipGetLastError();
            }
        }
        if (graphs_enabled && since_repack >= 1 && graph[parity]) {
            HIPCHK(hipGraphLaunch(graph[parity], s));
            replays++;
        } else {
            cell_iteration(m, s, pack, x0, xc, xn);
        }
        launches++;
        since_repack++;
        std::swap(xc, xn);
        parity ^= 1;
    }

    // Final energy violation and grid of every request active[idx[j]] (reads
    // xc): one slice table for all of them (with each slice's place in the
    // packing's table, for its reference energy), one violation launch over
    // it, one unpack launch over a job table, and the download queued on
    // s_copy; finish_retire() lands the results. The previous retirement
    // must have been landed (its results are in h_out).
    void retire(const std::vector<size_t>& idx) {
        hipStream_t s = m.stream;
        const size_t k = idx.size();
        std::vector<Slice> sl;
        std::vector<int> sidx;
        std::vector<CellJob> jobs(k);
        std::vector<size_t> out_off(k);
        size_t out_bytes = 0;
        int max_tok = 0;
        for (size_t j = 0; j < k; j++) {
            const Request& r = *active[idx[j]];
            const int n = r.dims.ny * r.dims.nx;
            for (int t = 0; t < r.dims.T; t++) {
                sl.push_back(Slice{r.base + t * n, n, (int)j});
                sidx.push_back(r.slice0 + t);  // slice0: this window's first slice in the packing's table
            }
            out_off[j] = out_bytes;
            jobs[j] = CellJob{(long)out_bytes, r.base, r.dims.tokens()};
            max_tok = std::max(max_tok, r.dims.tokens());
            out_bytes += aligned(12 * (size_t)r.dims.tokens());
        }
        reserve(d_out, cap_out, out_bytes);
        reserve(d_rsl, cap_rsl, sl.size());
        reserve(d_ridx, cap_ridx, sidx.size());
        reserve(d_jobs, cap_jobs, k);
        reserve(d_viol, cap_viol, k);
        // h_rtab, not h_tab: these uploads may still be queued when the
        // next repack writes h_tab
        const size_t o1 = aligned(sl.size() * sizeof(Slice)), o2 = o1 + aligned(sidx.size() * sizeof(int));
        reserve_host(h_rtab, cap_hrtab, o2 + k * sizeof(CellJob));
        std::memcpy(h_rtab, sl.data(), sl.size() * sizeof(Slice));
        std::memcpy(h_rtab + o1, sidx.data(), sidx.size() * sizeof(int));
        std::memcpy(h_rtab + o2, jobs.data(), k * sizeof(CellJob));
        HIPCHK(hipMemcpyAsync(d_rsl, h_rtab, sl.size() * sizeof(Slice), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_ridx, h_rtab + o1, sidx.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_jobs, h_rtab + o2, k * sizeof(CellJob), hipMemcpyHostToDevice, s));
        fill_u32(s, reinterpret_cast<unsigned*>(d_viol), (int)(2 * k), 0u);
        cell_violation(s, (int)sl.size(), d_rsl, xc, pack.eref_t, d_viol, m.cfg.g, m.cfg.norm_mean, m.cfg.norm_std,
                       m.cfg.norm_depth, m.cfg.energy_r_max, d_ridx);
        cell_unpack_jobs(s, (int)k, d_jobs, max_tok, xc, d_out);
        reserve_host(h_out, cap_hout, out_bytes + k * sizeof(double));
        HIPCHK(hipEventRecord(ev_unpacked, s));
        HIPCHK(hipStreamWaitEvent(s_copy, ev_unpacked, 0));
        HIPCHK(hipMemcpyAsync(h_out, d_out, out_bytes, hipMemcpyDeviceToHost, s_copy));
        HIPCHK(hipMemcpyAsync(h_out + out_bytes, d_viol, k * sizeof(double), hipMemcpyDeviceToHost, s_copy));
        HIPCHK(hipEventRecord(ev_out, s_copy));
        deferred_bytes = out_bytes;
        for (size_t j = 0; j < k; j++) {
            Request& r = *active[idx[j]];
            r.res.iters = r.iters;
            r.res.resid = r.resid;
            r.res.converged = r.resid < m.cfg.tol ? 1 : 0;
            landings.push_back(Landing{&r, out_off[j], r.dims.T, -1, r.iters, r.resid < m.cfg.tol, false});
        }
    }

    // The host half of a retirement: wait for its download, copy each
    // result into place (a request's grid; a rollout window's slices after
    // its first into the trajectory, with its time, violation and
    // iterations; the trajectory becomes the result after the last window)
    // and publish the requests that ended.
    void finish_retire() {
        if (landings.empty()) return;
        HIPCHK(hipEventSynchronize(ev_out));
        const size_t k = landings.size();
        std::vector<Copy> copies;
        for (const Landing& L : landings) {
            Request& r = *L.r;
            const size_t n = (size_t)r.dims.ny * r.dims.nx;
            const unsigned char* src = h_out + L.off;
            if (L.t0 < 0) {
                copies.push_back(Copy{r.grid.data(), src, 12 * (size_t)L.T * n});
                continue;
            }
            for (int c = 0; c < 3; c++)  // slice 0 of a window is the previous window's last
                copies.push_back(Copy{&r.traj[((size_t)c * r.ticks + L.t0 + 1) * n],
                                      src + ((size_t)c * L.T + 1) * n * sizeof(float), (size_t)(L.T - 1) * n * sizeof(float)});
        }
        copy_all(copies);
        std::vector<double> v(k);
        std::memcpy(v.data(), h_out + deferred_bytes, k * sizeof(double));
        const double t = now_s();
        for (size_t j = 0; j < k; j++) {
            const Landing& L = landings[j];
            Request& r = *L.r;
            if (L.t0 < 0) {
                r.res.violation = v[j];
                r.res.done_s = t;
                continue;
            }
            r.win_done.push_back(t);
            r.win_viol.push_back(v[j]);
            r.win_iters.push_back(L.iters);
            r.converged_all = r.converged_all && L.converged;
            if (!L.last) continue;
            // the result of a rollout is its trajectory; counters over all windows
            r.grid.swap(r.traj);
            r.dims.T = r.ticks;
            int it = 0;
            double vm = 0;
            for (size_t w = 0; w < r.win_iters.size(); w++) {
                it += r.win_iters[w];
                vm = std::max(vm, r.win_viol[w]);
            }
            r.res.iters = it;
            r.res.violation = vm;
            r.res.converged = r.converged_all ? 1 : 0;
            r.res.done_s = t;
        }
        landings.clear();
        std::lock_guard<std::mutex> lk(mu);
        for (auto& r : deferred) {
            const uint64_t id = r->id;
            done[id] = std::move(r);
        }
        deferred.clear();
        n_active = active.size();
        cv_done.notify_all();
    }

    // A rollout whose window has just retired moves to its next window
    // (made on the device from this one's last slice at the next repack);
    // false after its last window. Its results land in finish_retire().
    bool advance_window(Request& r) {
        r.wi++;
        if (r.wi >= window_count(r.ticks, r.window)) return false;
        r.prev_T = r.dims.T;
        r.dims.T = window_slices(r.ticks, r.window, r.wi);
        r.iters = 0;
        r.resid = 0.f;
        r.watch = StallWatch{};
        return true;
    }

    void loop() {
        DeviceGuard dg(m.device);
        HIPCHK(hipStreamCreateWithFlags(&s_copy, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&ev_unpacked, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ev_packed, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ev_uploaded, hipEventDisableTiming));
        for (;;) {
            bool changed = false;
            {
                bool work;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    work = stop || !pending.empty() || !active.empty();
                }
                if (!work) finish_retire();  // nothing to overlap it with
            }
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !pending.empty() || !active.empty(); });
                if (stop) return;
                size_t used = 0;
                for (auto& r : active) used += r->bytes;
                while (!pending.empty() && (int)active.size() < max_active) {
                    Request* r = pending.front().get();
                    if (used + r->bytes > budget) break;
                    used += r->bytes;
                    r->res.start_s = now_s();
                    active.push_back(std::move(pending.front()));
                    pending.pop_front();
                    changed = true;
                }
                n_active = active.size() + deferred.size();
                peak_active = std::max(peak_active, (int)active.size());
                peak_bytes = std::max(peak_bytes, used);
            }
            if (active.empty()) continue;
            const double t_a = now_s();
            if (changed || !packed) repack();
            const double t_b = now_s();
            run_iteration();
            HIPCHK(hipMemcpyAsync(h_resid, pack.resid, 4 * active.size(), hipMemcpyDeviceToHost, m.stream));
            prestage();       // beside the iteration
            finish_retire();  // the last turn's results, beside the iteration too
            // requests that arrive while the iteration runs are staged as they come
            for (;;) {
                const hipError_t q = hipStreamQuery(m.stream);
                if (q != hipErrorNotReady) {
                    HIPCHK(q);
                    break;
                }
                if (!prestage()) std::this_thread::yield();
            }
            HIPCHK(hipStreamSynchronize(m.stream));
            const double t_c = now_s();
            ph_turns += 1;
            ph_repack += t_b - t_a;
            ph_iter += t_c - t_b;
            std::vector<size_t> retiring;
            for (size_t i = 0; i < active.size(); i++) {
                Request& r = *active[i];
                r.iters++;
                std::memcpy(&r.resid, &h_resid[i], 4);
                const bool stalled = r.watch.update(r.resid) && m.cfg.stall_stop;
                if (r.resid < m.cfg.tol || r.iters >= m.cfg.max_iter || stalled) {
                    r.res.stalled = r.res.stalled || (r.resid >= m.cfg.tol && stalled) ? 1 : 0;
                    retiring.push_back(i);
                }
            }
            if (retiring.empty()) continue;
            retire(retiring);
            std::vector<bool> leaving(active.size(), false);
            for (size_t j = 0; j < retiring.size(); j++) {
                Request& r = *active[retiring[j]];
                Landing& L = landings[landings.size() - retiring.size() + j];
                if (r.ticks > 0) {
                    L.t0 = r.wi * (r.window - 1);
                    if (advance_window(r)) continue;  // stays active with its next window
                    L.last = true;
                }
                leaving[retiring[j]] = true;
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                std::vector<std::unique_ptr<Request>> keep;
                for (size_t i = 0; i < active.size(); i++) {
                    if (leaving[i]) deferred.push_back(std::move(active[i]));
                    else keep.push_back(std::move(active[i]));
                }
                active = std::move(keep);
                n_active = active.size() + deferred.size();
            }
            packed = false;  // survivors move to a compact layout next turn
            ph_finish += now_s() - t_c;
        }
    }

    void start() {
        worker = std::thread([this] {
            try {
                loop();
            } catch (const Error& e) {
                std::lock_guard<std::mutex> lk(mu);
                worker_error = e.what();
                worker_code = e.code;
            } catch (const std::exception& e) {
                std::lock_guard<std::mutex> lk(mu);
                worker_error = e.what();
                worker_code = J2K_ERR_POLL;
            }
            cv_done.notify_all();
        });
    }
};

}  // namespace j2k

using namespace j2k;

static Server& server_of(void* h) {
    if (!h) fail(J2K_ERR_INVALID, "server handle is NULL");
    return *static_cast<Server*>(h);
}

extern "C" {

int j2k_server_create(void* model, size_t budget, int max_active, void** out) {
    return guard([&] {
        if (!model) fail(J2K_ERR_INVALID, "model handle is NULL");
        if (!out) fail(J2K_ERR_INVALID, "out_server is NULL");
        Model& m = *static_cast<Model*>(model);
        if (m.cfg.form != J2K_FORM_CELL_ST) fail(J2K_ERR_INVALID, "the server runs cell_st models");
        if (max_active <= 0) fail(J2K_ERR_INVALID, "max_active must be positive");
        if (m.tune.enabled) {
            // tune before the worker starts, up to the largest pack the budget admits
            const size_t cap = budget / bytes_per_token(m);
            autotune(m, (long)std::min<size_t>(cap, size_t(1) << 20));
        }
        auto s = std::make_unique<Server>(m, budget, max_active);
        {  // pinned staging for the largest pack the budget admits, up to 2^24 tokens
            DeviceGuard dg(m.device);
            const size_t tokens = std::min<size_t>(budget / bytes_per_token(m), size_t(1) << 24);
            const size_t slack = 256 * (size_t)std::min(max_active, 4096);
            s->reserve_host(s->h_stage[0], s->cap_stage[0], 13 * tokens + 2 * slack);
            s->reserve_host(s->h_stage[1], s->cap_stage[1], 13 * tokens + 2 * slack);
            s->reserve_host(s->h_out, s->cap_hout, 12 * tokens + 2 * slack);
            // the turn's tables: about 1.2 bytes a token (attention tiles, cell tiles) and a few per request
            s->reserve_host(s->h_tab, s->cap_htab, 2 * tokens + 64 * slack);
            s->reserve(s->d_jobs, s->cap_jobs, (size_t)std::min(max_active, 1 << 16));
        }
        s->start();
        *out = s.release();
    });
}

int j2k_server_submit(void* h, const float* grid, const uint8_t* known, const float* meta, int T, int ny, int nx,
                      uint64_t* out_id) {
    return guard([&] {
        Server& s = server_of(h);
        if (!grid || !known || !meta) fail(J2K_ERR_INVALID, "grid, known and meta must be non-NULL");
        CellDims d{T, ny, nx};
        check_cell_dims(s.m, d);
        auto r = std::make_unique<Request>();
        r->dims = d;
        const size_t M = (size_t)d.tokens();
        r->grid.resize(3 * M);
        r->known.assign(known, known + M);
        r->meta.assign(meta, meta + s.m.cfg.meta_dim);
        const int frc = j2k_fill_blank(grid, known, T, ny, nx, r->grid.data());
        if (frc) fail(frc, "blank fill failed");
        r->bytes = M * bytes_per_token(s.m);
        if (r->bytes > s.budget) fail(J2K_ERR_OOM, "request needs more device memory than the whole budget");
        r->res.submit_s = now_s();
        std::lock_guard<std::mutex> lk(s.mu);
        if (s.worker_code) fail(s.worker_code, "server stopped: " + s.worker_error);
        r->id = r->res.id = s.next_id++;
        if (out_id) *out_id = r->id;
        s.pending.push_back(std::move(r));
        s.cv_work.notify_one();
    });
}

int j2k_server_submit_rollout(void* h, const float* state, const float* meta, int ny, int nx, int window, int ticks,
                              uint64_t* out_id) {
    return guard([&] {
        Server& s = server_of(h);
        if (!state || !meta) fail(J2K_ERR_INVALID, "state and meta must be non-NULL");
        if (window < 2) fail(J2K_ERR_INVALID, "a rollout window needs at least 2 slices");
        if (ticks < 2) fail(J2K_ERR_INVALID, "a rollout needs at least 2 ticks");
        CellDims d{window, ny, nx};
        check_cell_dims(s.m, d);
        if (window_slices(ticks, window, window_count(ticks, window) - 1) < 2)
            fail(J2K_ERR_INVALID, "internal: last rollout window shorter than 2 slices");
        auto r = std::make_unique<Request>();
        r->dims = d;
        r->ticks = ticks;
        r->window = window;
        r->meta.assign(meta, meta + s.m.cfg.meta_dim);
        const size_t n = (size_t)ny * nx;
        r->traj.assign(3 * (size_t)ticks * n, 0.f);
        for (int c = 0; c < 3; c++) std::memcpy(&r->traj[(size_t)c * ticks * n], state + c * n, n * sizeof(float));
        set_window(*r, state, window_slices(ticks, window, 0));
        r->bytes = (size_t)d.tokens() * bytes_per_token(s.m);
        if (r->bytes > s.budget) fail(J2K_ERR_OOM, "a window needs more device memory than the whole budget");
        r->res.submit_s = now_s();
        std::lock_guard<std::mutex> lk(s.mu);
        if (s.worker_code) fail(s.worker_code, "server stopped: " + s.worker_error);
        r->id = r->res.id = s.next_id++;
        if (out_id) *out_id = r->id;
        s.pending.push_back(std::move(r));
        s.cv_work.notify_one();
    });
}

int j2k_server_windows(void* h, uint64_t id, int n, int* count, double* latency_s, double* violation, int* iters) {
    return guard([&] {
        Server& s = server_of(h);
        std::lock_guard<std::mutex> lk(s.mu);
        auto it = s.done.find(id);
        if (it == s.done.end()) {
            if (s.worker_code) fail(s.worker_code, "server stopped: " + s.worker_error);
            if (id == 0 || id >= s.next_id) fail(J2K_ERR_INVALID, "unknown request id");
            throw Error(J2K_PENDING, "request still running");
        }
        const Request& r = *it->second;
        if (r.ticks == 0) fail(J2K_ERR_INVALID, "not a rollout request");
        const int w = (int)r.win_done.size();
        if (count) *count = w;
        for (int i = 0; i < std::min(n, w); i++) {
            const double start = i == 0 ? r.res.submit_s : r.win_done[i - 1];
            if (latency_s) latency_s[i] = r.win_done[i] - start;
            if (violation) violation[i] = r.win_viol[i];
            if (iters) iters[i] = r.win_iters[i];
        }
    });
}

int j2k_server_poll(void* h, uint64_t id, j2k_result* out, float* grid_out) {
    return guard([&] {
        Server& s = server_of(h);
        std::lock_guard<std::mutex> lk(s.mu);
        auto it = s.done.find(id);
        if (it == s.done.end()) {
            if (s.worker_code) fail(s.worker_code, "server stopped: " + s.worker_error);
            if (id == 0 || id >= s.next_id) fail(J2K_ERR_INVALID, "unknown request id");
            throw Error(J2K_PENDING, "request still running");
        }
        if (out) *out = it->second->res;
        if (grid_out) std::memcpy(grid_out, it->second->grid.data(), it->second->grid.size() * sizeof(float));
    });
}

int j2k_server_drain(void* h) {
    return guard([&] {
        Server& s = server_of(h);
        std::unique_lock<std::mutex> lk(s.mu);
        s.cv_done.wait(lk, [&] { return s.worker_code != 0 || (s.pending.empty() && s.n_active == 0); });
        if (s.worker_code) fail(s.worker_code, "server stopped: " + s.worker_error);
    });
}

int j2k_server_stats(void* h, uint64_t* launches, uint64_t* replays, int* peak_active, size_t* peak_bytes) {
    return guard([&] {
        Server& s = server_of(h);
        std::lock_guard<std::mutex> lk(s.mu);
        if (launches) *launches = s.launches;
        if (replays) *replays = s.replays;
        if (peak_active) *peak_active = s.peak_active;
        if (peak_bytes) *peak_bytes = s.peak_bytes;
    });
}

int j2k_server_destroy(void* h) {
    return guard([&] { delete static_cast<Server*>(h); });
}

}  // extern "C"
