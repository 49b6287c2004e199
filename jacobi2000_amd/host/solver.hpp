// C++ mirror of the reference `solvers` crate's shallow-water API
// (crates/solvers/src/shallow_water.rs:16-244) over libj2k's device solver:
// Params, State (flat, mass, add_drop) and ShallowWater (new, step, steps).
// Every step runs on the GPU, bit-identical to the reference's f32 step.
// Header-only; link with -lj2k.
#pragma once

#include <cmath>
#include <memory>
#include <vector>

#include "engine.hpp"

namespace solvers {

struct Params {
    int nx = 64, ny = 64;
    float dx = 1, dt = 0, g = 9.81f, depth = 1, nu = 0;
    int cells() const { return nx * ny; }
    j2k_sw_params c() const { return j2k_sw_params{nx, ny, dx, dt, g, depth, nu}; }
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
    // shallow_water.rs:96-98: sum of h in f64, left to right
    double mass() const {
        double m = 0;
        for (float x : h) m += (double)x;
        return m;
    }
    // shallow_water.rs:102-111: the bump as the reference computes it in f32
    void add_drop(const Params& p, float cx, float cy, float amp, float sigma) {
        std::vector<float> f(p.cells());
        const j2k_sw_params c = p.c();
        engine::check(j2k_sw_drop_field(&c, cx, cy, amp, sigma, f.data()));
        for (size_t i = 0; i < f.size(); i++) h[i] += f[i];
    }
};

// ShallowWater::new (shallow_water.rs:136-147) on `device`.
class ShallowWater {
   public:
    explicit ShallowWater(Params p, int device = 0) : p_(p) {
        const j2k_sw_params c = p.c();
        void* h = nullptr;
        engine::check(j2k_sw_create(&c, device, &h));
        h_.reset(h, Del{});
    }
    // shallow_water.rs:149-238
    void step(State& s) { steps(s, 1); }
    // shallow_water.rs:240-244
    void steps(State& s, int n) { engine::check(j2k_sw_steps(h_.get(), s.h.data(), s.u.data(), s.v.data(), 1, n)); }
    const Params& params() const { return p_; }
    void* handle() const { return h_.get(); }

   private:
    struct Del {
        void operator()(void* p) const { j2k_sw_destroy(p); }
    };
    Params p_;
    std::shared_ptr<void> h_;
};

}  // namespace solvers
