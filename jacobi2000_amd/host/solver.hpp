// Host restatement of the reference shallow-water solver
// (crates/solvers/src/shallow_water.rs:16-244): the rollout CLI's truth
// partner and initial conditions. Forward-backward scheme on a collocated
// grid with reflecting walls through ghost-cell symmetry.
#pragma once

#include <cmath>
#include <vector>

namespace solvers {

struct Params {
    int nx = 64, ny = 64;
    float dx = 1, dt = 0, g = 9.81f, depth = 1, nu = 0;
    int cells() const { return nx * ny; }
};

struct State {
    std::vector<float> h, u, v;
    // shallow_water.rs:88-94
    static State flat(const Params& p) {
        State s;
        s.h.assign(p.cells(), p.depth);
        s.u.assign(p.cells(), 0.f);
        s.v.assign(p.cells(), 0.f);
        return s;
    }
    // shallow_water.rs:102-111
    void add_drop(const Params& p, float cx, float cy, float amp, float sigma) {
        const float inv = 1.0f / (2.0f * sigma * sigma);
        for (int y = 0; y < p.ny; y++)
            for (int x = 0; x < p.nx; x++) {
                const float dx = (float)x - cx, dy = (float)y - cy;
                h[y * p.nx + x] += amp * std::exp(-(dx * dx + dy * dy) * inv);
            }
    }
};

class ShallowWater {
   public:
    explicit ShallowWater(Params p)
        : p_(p), fx_((p.nx + 1) * p.ny), fy_(p.nx * (p.ny + 1)), hn_(p.cells()), un_(p.cells()), vn_(p.cells()) {}

    // shallow_water.rs:149-238
    void step(State& s) {
        const int nx = p_.nx, ny = p_.ny;
        auto idx = [&](int x, int y) { return y * nx + x; };
        auto mirror = [](int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); };
        for (int y = 0; y < ny; y++) {
            for (int x = 1; x < nx; x++) {
                const int a = idx(x - 1, y), b = idx(x, y);
                fx_[y * (nx + 1) + x] = 0.5f * (s.h[a] + s.h[b]) * 0.5f * (s.u[a] + s.u[b]);
            }
            fx_[y * (nx + 1)] = 0.f;
            fx_[y * (nx + 1) + nx] = 0.f;
        }
        for (int x = 0; x < nx; x++) {
            for (int y = 1; y < ny; y++) {
                const int a = idx(x, y - 1), b = idx(x, y);
                fy_[y * nx + x] = 0.5f * (s.h[a] + s.h[b]) * 0.5f * (s.v[a] + s.v[b]);
            }
            fy_[x] = 0.f;
            fy_[ny * nx + x] = 0.f;
        }
        const float r = p_.dt / p_.dx;
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const float dfx = fx_[y * (nx + 1) + x + 1] - fx_[y * (nx + 1) + x];
                const float dfy = fy_[(y + 1) * nx + x] - fy_[y * nx + x];
                hn_[idx(x, y)] = s.h[idx(x, y)] - r * (dfx + dfy);
            }
        const float inv2dx = 1.0f / (2.0f * p_.dx), invdx2 = 1.0f / (p_.dx * p_.dx);
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const int i = idx(x, y);
                const float sxm = x == 0 ? -1.f : 1.f, sxp = x + 1 == nx ? -1.f : 1.f;
                const float sym = y == 0 ? -1.f : 1.f, syp = y + 1 == ny ? -1.f : 1.f;
                const float hl = hn_[idx(mirror(x - 1, nx), y)], hr = hn_[idx(mirror(x + 1, nx), y)];
                const float hd = hn_[idx(x, mirror(y - 1, ny))], hu = hn_[idx(x, mirror(y + 1, ny))];
                const float ul = sxm * s.u[idx(mirror(x - 1, nx), y)], ur = sxp * s.u[idx(mirror(x + 1, nx), y)];
                const float ud = s.u[idx(x, mirror(y - 1, ny))], uu = s.u[idx(x, mirror(y + 1, ny))];
                const float vl = s.v[idx(mirror(x - 1, nx), y)], vr = s.v[idx(mirror(x + 1, nx), y)];
                const float vd = sym * s.v[idx(x, mirror(y - 1, ny))], vu = syp * s.v[idx(x, mirror(y + 1, ny))];
                const float u0 = s.u[i], v0 = s.v[i];
                const float dudx = (ur - ul) * inv2dx, dudy = (uu - ud) * inv2dx;
                const float dvdx = (vr - vl) * inv2dx, dvdy = (vu - vd) * inv2dx;
                const float dhdx = (hr - hl) * inv2dx, dhdy = (hu - hd) * inv2dx;
                const float lap_u = (ul + ur + ud + uu - 4.0f * u0) * invdx2;
                const float lap_v = (vl + vr + vd + vu - 4.0f * v0) * invdx2;
                un_[i] = u0 + p_.dt * (-u0 * dudx - v0 * dudy - p_.g * dhdx + p_.nu * lap_u);
                vn_[i] = v0 + p_.dt * (-u0 * dvdx - v0 * dvdy - p_.g * dhdy + p_.nu * lap_v);
            }
        s.h = hn_;
        s.u = un_;
        s.v = vn_;
    }

   private:
    Params p_;
    std::vector<float> fx_, fy_, hn_, un_, vn_;
};

// crates/engine/src/metrics.rs:9-23: Pearson correlation, f64 accumulation,
// 0 for a constant field.
inline float correlation(const std::vector<float>& a, const std::vector<float>& b) {
    const double n = (double)a.size();
    double ma = 0, mb = 0;
    for (size_t i = 0; i < a.size(); i++) {
        ma += a[i];
        mb += b[i];
    }
    ma /= n;
    mb /= n;
    double sab = 0, saa = 0, sbb = 0;
    for (size_t i = 0; i < a.size(); i++) {
        const double dx = a[i] - ma, dy = b[i] - mb;
        sab += dx * dy;
        saa += dx * dx;
        sbb += dy * dy;
    }
    if (saa <= 0 || sbb <= 0) return 0.f;
    return (float)(sab / std::sqrt(saa * sbb));
}

}  // namespace solvers
