// The cell_st fixed-point iteration over a packed set of requests.
#include "cell.h"

#include <string>

namespace j2k {

void CellTables::build(const std::vector<CellDims>& reqs, int tpb, bool per_token_req, bool generic_tiles) {
    segs.clear();
    slices.clear();
    tok_req.clear();
    M = 0;
    for (size_t r = 0; r < reqs.size(); r++) {
        const CellDims& d = reqs[r];
        segs.push_back(Seg{M, d.T, d.ny, d.nx});
        const int n = d.ny * d.nx;
        for (int t = 0; t < d.T; t++) slices.push_back(Slice{M + t * n, n, (int)r});
        if (per_token_req) tok_req.insert(tok_req.end(), (size_t)d.tokens(), (int)r);
        M += d.tokens();
    }
    tiles.clear();
    if (generic_tiles) build_tiles(segs, tpb, tiles);
    build_cell_tiles(segs, ctiles);
}

bool cell_needs_generic_tiles(const Model& m) {
    const AttnGeo g = m.geo();
    const int elem = m.bf16() ? 2 : 4;  // the activations the attention reads (f32 in every other form)
    return !(g.kind == ATTN_CELL_ST && attn_cell_ok(m.cfg.dim / m.cfg.heads, g.radius, g.time_radius, elem, KV_JOINT));
}

void check_cell_dims(const Model& m, const CellDims& d) {
    if (m.cfg.form != J2K_FORM_CELL_ST) fail(J2K_ERR_INVALID, "this entry point needs a cell_st model");
    if (d.T <= 0 || d.ny <= 0 || d.nx <= 0) fail(J2K_ERR_SHAPE, "grid dimensions must be positive");
    const int w = 2 * m.cfg.radius + 1;
    if (d.ny < w || d.nx < w)
        fail(J2K_ERR_SHAPE, "grid " + std::to_string(d.ny) + "x" + std::to_string(d.nx) +
                                " is smaller than the periodic stencil width " + std::to_string(w));
    if ((long)d.T * d.ny * d.nx > (1L << 28)) fail(J2K_ERR_SHAPE, "grid too large (over 2^28 cells)");
}

// The trunk over the pack, ending in the readout `ro`; clears zero[0 ..
// n_zero) on the way.
static void embed_and_trunk(Model& m, hipStream_t s, const CellPack& p, const float* x4, float* X, const Readout& ro,
                            unsigned* zero = nullptr, int n_zero = 0) {
    const int d = m.cfg.dim;
    const LayerW& L0 = m.L.at(0);
    if ((m.bf16() || m.split()) && rows_fused_ok(d, m.cfg.ff_inner) && L0.wqkv_f) {
        // the embedding, LN1 and the first qkv product in one launch
        // the first row-panel launch remakes the embedding rather than read X
        const bool emb_in = rows_embed_in_trunk(m, p.M);
        RowsArgs a;
        a.X = emb_in ? nullptr : X;
        a.n1g = L0.n1g;
        a.n1b = L0.n1b;
        a.Wqkv = L0.wqkv_f;
        a.bqkv = L0.bqkv;
        a.QKV = m.get<char>(S_QKV, (size_t)p.M * 3 * d * (m.split() ? 4 : 2));
        a.zero = zero;
        a.n_zero = p.M > 0 ? n_zero : 0;
        if (p.M <= 0 && n_zero > 0) fill_u32(s, zero, n_zero, 0u);
        {
            TScope t(m, s, KC_GEMM, 2.0 * p.M * d * 3 * d);
            rows_embed_qkv(s, p.M, d, m.split(), x4, m.embed_w, m.embed_b, p.mvec, p.tok_req, a);
        }
        RowsEmbed e;
        e.x4 = x4;
        e.we = m.embed_w;
        e.be = m.embed_b;
        e.mvec = p.mvec;
        e.tok_req = p.tok_req;
        if (!trunk(m, s, p.M, X, TRUNK_QKV, plan_of(p), &ro, emb_in ? &e : nullptr))
            cell_readout(s, p.M, d, X, ro.g, ro.b, ro.w, ro.bias, ro.x0, ro.xn, ro.y);
        return;
    }
    if (n_zero > 0) fill_u32(s, zero, n_zero, 0u);
    if ((m.bf16() || m.split()) && mlp_fused_ok(d, m.cfg.ff_inner) && L0.wqkv_f && L0.w1_f &&
        m.tune.choice(TO_ROWS, p.M) == 2) {
        // d = 256 with the whole block in one launch: the embedding, LN1 and
        // the first qkv product in one launch too
        RowsArgs a;
        a.X = X;
        a.n1g = L0.n1g;
        a.n1b = L0.n1b;
        a.Wqkv = L0.wqkv_f;
        a.bqkv = L0.bqkv;
        a.QKV = m.get<char>(S_QKV, (size_t)p.M * 3 * d * (m.split() ? 4 : 2));
        {
            TScope t(m, s, KC_GEMM, 2.0 * p.M * d * 3 * d);
            rows256_embed_qkv(s, p.M, d, m.split(), x4, m.embed_w, m.embed_b, p.mvec, p.tok_req, a);
        }
        if (!trunk(m, s, p.M, X, TRUNK_QKV, plan_of(p), &ro))
            cell_readout(s, p.M, d, X, ro.g, ro.b, ro.w, ro.bias, ro.x0, ro.xn, ro.y);
        return;
    }
    if (m.bf16()) {
        uint16_t* Hb = m.get<uint16_t>(S_HB, (size_t)p.M * d);
        cell_embed_ln(s, p.M, d, x4, m.embed_w, m.embed_b, p.mvec, p.tok_req, L0.n1g, L0.n1b, X, nullptr, Hb);
    } else {
        float* H = m.get<float>(S_H, (size_t)p.M * d);
        cell_embed_ln(s, p.M, d, x4, m.embed_w, m.embed_b, p.mvec, p.tok_req, L0.n1g, L0.n1b, X, H, nullptr);
    }
    if (!trunk(m, s, p.M, X, TRUNK_H, plan_of(p), &ro))
        cell_readout(s, p.M, d, X, ro.g, ro.b, ro.w, ro.bias, ro.x0, ro.xn, ro.y);
}

// The tables of one request depend on its grid only: built and uploaded on
// the first call with a grid, reused while the next calls keep it. With
// d_meta, also the request's meta vector (else the caller launches it).
void cell_setup(Model& m, hipStream_t s, const CellDims& d, const float* d_meta, CellRun& run) {
    check_cell_dims(m, d);
    CellPack& p = run.pack;
    auto& c = m.cell_tables;
    m.p_tables_batch = -1;  // the patch2d tables share these slots
    if (c.T != d.T || c.ny != d.ny || c.nx != d.nx) {
        c.T = -1;
        run.tab.build({d}, attn_tokens_per_block(m.cfg.dim), false);
        HIPCHK(hipStreamSynchronize(s));
        Seg* segs = m.get<Seg>(S_SEGS, run.tab.segs.size());
        Tile* tiles = m.get<Tile>(S_TILES, run.tab.tiles.size());
        Slice* slices = m.get<Slice>(S_SLICES, run.tab.slices.size());
        CellTile* ctiles = m.get<CellTile>(S_CTILES, run.tab.ctiles.size());
        HIPCHK(hipMemcpy(ctiles, run.tab.ctiles.data(), run.tab.ctiles.size() * sizeof(CellTile), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(segs, run.tab.segs.data(), run.tab.segs.size() * sizeof(Seg), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(tiles, run.tab.tiles.data(), run.tab.tiles.size() * sizeof(Tile), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(slices, run.tab.slices.data(), run.tab.slices.size() * sizeof(Slice), hipMemcpyHostToDevice));
        c.M = run.tab.M;
        c.n_slices = (int)run.tab.slices.size();
        c.n_tiles = (int)run.tab.tiles.size();
        c.n_ctiles = (int)run.tab.ctiles.size();
        c.T = d.T;
        c.ny = d.ny;
        c.nx = d.nx;
    }
    p.M = c.M;
    p.n_req = 1;
    p.n_slices = c.n_slices;
    p.n_tiles = c.n_tiles;
    p.n_ctiles = c.n_ctiles;
    p.segs = static_cast<Seg*>(m.current(S_SEGS));
    p.tiles = static_cast<Tile*>(m.current(S_TILES));
    p.slices = static_cast<Slice*>(m.current(S_SLICES));
    p.ctiles = static_cast<CellTile*>(m.current(S_CTILES));
    p.tok_req = nullptr;
    p.mvec = m.get<float>(S_MVEC, m.cfg.dim);
    if (d_meta) meta_linear(s, m.cfg.meta_dim, m.cfg.dim, d_meta, m.meta_w, m.meta_b, p.mvec);
    double* targ = m.get<double>(S_TARG, 2 * (size_t)p.n_slices);
    p.mass_t = targ;
    p.eref_t = targ + p.n_slices;
    p.resid = m.get<unsigned>(S_RESID, 4);
}

void cell_iteration(Model& m, hipStream_t s, const CellPack& p, const float* x0, const float* xc, float* xn,
                    float* grid) {
    const j2k_config& c = m.cfg;
    float* X = m.get<float>(S_X, (size_t)p.M * c.dim);
    Readout ro;
    ro.g = m.norm_g;
    ro.b = m.norm_b;
    ro.w = m.ro_w;
    ro.bias = m.ro_b;
    ro.x0 = x0;
    ro.xn = xn;
    // the residual words (and the violation after a single request's word)
    // start at zero for the projection's maxima
    embed_and_trunk(m, s, p, xc, X, ro, p.resid, p.viol ? 4 : p.n_req);
    // read xn and xc, write xn: three float4 per cell (and three floats of grid)
    TScope t(m, s, KC_PROJ, (grid ? 60.0 : 48.0) * p.M);
    cell_project(s, p.n_slices, p.slices, xn, xc, p.mass_t, p.eref_t, p.resid, c.g, c.norm_mean, c.norm_std,
                 c.norm_depth, c.energy_r_max, p.viol, grid, p.M);
}

void cell_model(Model& m, hipStream_t s, const CellPack& p, const float* x4, float* y) {
    const j2k_config& c = m.cfg;
    float* X = m.get<float>(S_X, (size_t)p.M * c.dim);
    Readout ro;
    ro.g = m.norm_g;
    ro.b = m.norm_b;
    ro.w = m.ro_w;
    ro.bias = m.ro_b;
    ro.y = y;
    embed_and_trunk(m, s, p, x4, X, ro);
}

}  // namespace j2k
