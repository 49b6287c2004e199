// The cell_st fixed-point iteration over a packed set of requests.
#pragma once

#include <vector>

#include "model.h"

namespace j2k {

// Host description of one request's grid.
struct CellDims {
    int T, ny, nx;
    int tokens() const { return T * ny * nx; }
};

// Device tables of a packed set of requests (request r owns tokens
// [segs[r].base, segs[r].base + T*ny*nx) of every packed array).
struct CellPack {
    int M = 0, n_req = 0, n_slices = 0, n_tiles = 0;
    Seg* segs = nullptr;
    Tile* tiles = nullptr;
    CellTile* ctiles = nullptr;
    int n_ctiles = 0;
    Slice* slices = nullptr;
    int* tok_req = nullptr;  // per token request index (nullptr for one request)
    float* mvec = nullptr;   // [n_req, dim] meta embedding per request
    double* mass_t = nullptr;
    double* eref_t = nullptr;
    unsigned* resid = nullptr;  // [n_req] float bits of max |dx|
    // one request only: the iteration also leaves the energy violation of
    // its result here (double bits), zeroed with the residual
    unsigned long long* viol = nullptr;
};

// Host-side tables for a list of requests (packed in order).
struct CellTables {
    std::vector<Seg> segs;
    std::vector<Tile> tiles;
    std::vector<CellTile> ctiles;
    std::vector<Slice> slices;
    std::vector<int> tok_req;
    int M = 0;
    // generic_tiles false: leave `tiles` empty (the tiled cell attention,
    // which reads ctiles, takes every launch)
    void build(const std::vector<CellDims>& reqs, int tpb, bool per_token_req, bool generic_tiles = true);
};

// Whether the model's attention can fall back to the generic gather (and so
// needs its token tiles): false when the tiled cell attention exists for the
// model's head size and stencil in its default staging mode.
bool cell_needs_generic_tiles(const Model& m);

void check_cell_dims(const Model& m, const CellDims& d);

// One request's tables and packed-pack view.
struct CellRun {
    CellTables tab;
    CellPack pack;
};

// The tables of one request depend on its grid only: built and uploaded on
// the first call with a grid (synchronising the stream), reused while the
// next calls keep it (m.cell_tables.T = -1 forces a rebuild). With d_meta,
// also the request's meta vector.
void cell_setup(Model& m, hipStream_t s, const CellDims& d, const float* d_meta, CellRun& run);

// Stall rule: when the residual fails to drop below 0.99x the best value
// seen so far on two consecutive iterations, the iteration has reached the
// noise floor of its arithmetic; it is retired there (reported as stalled,
// not converged). A contracting sequence never trips it.
struct StallWatch {
    float best = 3.4e38f;
    int flat = 0;
    bool update(float r) {
        flat = r < 0.99f * best ? 0 : flat + 1;
        best = r < best ? r : best;
        return flat >= 2;
    }
};

inline AttnPlan plan_of(const CellPack& p) {
    AttnPlan a;
    a.segs = p.segs;
    a.tiles = p.tiles;
    a.n_tiles = p.n_tiles;
    a.ctiles = p.ctiles;
    a.n_ctiles = p.n_ctiles;
    return a;
}

// One iteration: xn = P(clamp_known(x0 + f(xc))), resid[r] = max|xn - xc| over blank cells.
void cell_iteration(Model& m, hipStream_t s, const CellPack& p, const float* x0, const float* xc, float* xn,
                    float* grid = nullptr);
// y = f(x) only (no clamp, no projection): y [M, 3].
void cell_model(Model& m, hipStream_t s, const CellPack& p, const float* x4, float* y);

}  // namespace j2k
