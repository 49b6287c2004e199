// C++ mirror of the reference `engine` crate API over libj2k's C ABI
// (crates/engine/src/lib.rs:30-68, gpu.rs:199-437, project.rs, norm.rs,
// solver_meta.rs). Header-only; link with -lj2k.
//
// Error behaviour: every failure throws engine::Error carrying the C-ABI
// status. Status 1-5 are the reference's lean_runtime::Error variants
// (NoAdapter, NoDevice, Limit, Map, Poll); 6-9 are the loader conditions on
// which the reference panics (bad config, missing tensor, non-F32, shape).
#pragma once

#include <array>
#include <cmath>
#include <fstream>
#include <iterator>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/j2k.h"

namespace engine {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != J2K_OK) throw Error(rc, j2k_last_error());
}

// crates/engine/src/norm.rs:13-34
struct Normalizer {
    std::array<float, 3> mean{}, std{};
    float depth = 0;
    // Pack a solver state (h, u, v planes of n cells) into a normalized [3, n] slice.
    void encode(const std::vector<float>& h, const std::vector<float>& u, const std::vector<float>& v,
                std::vector<float>& out) const {
        const size_t n = h.size();
        out.resize(3 * n);
        for (size_t i = 0; i < n; i++) {
            out[i] = (h[i] - depth - mean[0]) / std[0];
            out[n + i] = (u[i] - mean[1]) / std[1];
            out[2 * n + i] = (v[i] - mean[2]) / std[2];
        }
    }
};

constexpr size_t META_DIM = 5;

// crates/engine/src/solver_meta.rs:9-11
inline std::array<float, META_DIM> meta_of(float dx, float dt, float g, float depth, float nu) {
    return {std::log(dx), std::log(dt), std::log(g), std::log(depth), std::log(nu + 1e-6f)};
}

// Model + GpuModel in one: the checkpoint lives on the GPU.
class Model {
   public:
    static Model load(const std::string& dir, int precision = J2K_PREC_F32, int device = 0) {
        return from_bytes(read_file(dir + "/model.safetensors"), read_text(dir + "/config.json"), precision, device);
    }
    static Model from_bytes(const std::vector<uint8_t>& st, const std::string& config_json,
                            int precision = J2K_PREC_F32, int device = 0) {
        void* h = nullptr;
        check(j2k_model_create(st.data(), st.size(), config_json.c_str(), device, precision, &h));
        return Model(h);
    }

    const j2k_config& cfg() const { return cfg_; }
    Normalizer norm() const {
        Normalizer n;
        for (int i = 0; i < 3; i++) {
            n.mean[i] = cfg_.norm_mean[i];
            n.std[i] = cfg_.norm_std[i];
        }
        n.depth = cfg_.norm_depth;
        return n;
    }
    std::array<float, META_DIM> meta() const { return meta_of(cfg_.dx, cfg_.dt, cfg_.g, cfg_.depth, cfg_.nu); }

    // lib.rs:53-55: the increment, same shape as state [channels, ny, nx].
    std::vector<float> forward(const std::vector<float>& state, const float* meta) const {
        std::vector<float> out(state.size());
        check(j2k_forward(h_.get(), state.data(), meta, out.data()));
        return out;
    }
    // lib.rs:58-60: state + forward(state).
    std::vector<float> step(const std::vector<float>& state, const float* meta) const {
        std::vector<float> out(state.size());
        check(j2k_step(h_.get(), state.data(), meta, out.data()));
        return out;
    }
    // The projected rollout loop on the device; returns the energy trace.
    std::vector<double> rollout(std::vector<float>& state, const float* meta, int steps, bool proj_mass,
                                bool proj_energy, double* max_violation = nullptr) const {
        std::vector<double> e((size_t)steps + 1);
        check(j2k_rollout(h_.get(), state.data(), meta, steps, proj_mass, proj_energy, e.data(), max_violation));
        return e;
    }
    void* handle() const { return h_.get(); }

   private:
    struct Del {
        void operator()(void* p) const { j2k_model_destroy(p); }
    };
    std::shared_ptr<void> h_;
    j2k_config cfg_{};
    explicit Model(void* h) : h_(h, Del{}) { check(j2k_model_config(h, &cfg_)); }
    static std::vector<uint8_t> read_file(const std::string& p) {
        std::ifstream f(p, std::ios::binary);
        if (!f) throw Error(J2K_ERR_INVALID, "reading " + p);
        return std::vector<uint8_t>(std::istreambuf_iterator<char>(f), {});
    }
    static std::string read_text(const std::string& p) {
        std::vector<uint8_t> b = read_file(p);
        return std::string(b.begin(), b.end());
    }
};

// crates/engine/src/project.rs:9-54, run on the model's device (sums in f64).
namespace project {
inline double mass(const Model& m, const std::vector<float>& x, int n) {
    double out = 0;
    check(j2k_mass(m.handle(), x.data(), n, &out));
    return out;
}
inline double energy(const Model& m, const std::vector<float>& x, int n) {
    double out = 0;
    check(j2k_energy(m.handle(), x.data(), n, &out));
    return out;
}
inline void project_mass(const Model& m, std::vector<float>& x, float target, int n) {
    check(j2k_project_mass(m.handle(), x.data(), n, target));
}
inline void project_energy(const Model& m, std::vector<float>& x, float prev_energy, int n) {
    check(j2k_project_energy(m.handle(), x.data(), n, prev_energy));
}
}  // namespace project

}  // namespace engine
