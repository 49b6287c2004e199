// `rollout`: the reference CLI (crates/engine-cli/src/main.rs) over libj2k.
//
//   rollout --model <checkpoint-dir> [--steps N] [--backend auto|webgpu|hip]
//           [--case drop|dam_break] [--out field.npy] [--precision f32|bf16]
//
// Steps the shallow-water solver beside the model from the same initial
// condition and prints the model's correlation against it every step
// (main.rs:99-142). The whole loop runs on the GPU (j2k_rollout_truth):
// the solver, the model step with its mass projection and the correlation,
// with one copy in and one out; the lines print when it returns. The
// reference's backends are auto | cpu | webgpu (crates/engine-cli/src/
// backend.rs:32-66): here `webgpu` (the reference's GPU engine) and `hip`
// pick the HIP engine, `auto` takes it too, and a missing GPU exits 1 as the
// reference does when no adapter is granted. The CPU engine is the
// reference's own, so `--backend cpu` is refused with exit code 2.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

#include "engine.hpp"
#include "solver.hpp"

namespace {

struct Args {
    std::string model, backend = "auto", kase = "drop", out, precision = "f32";
    int steps = 50;
};

[[noreturn]] void usage() {
    std::fprintf(stderr,
                 "usage: rollout --model <checkpoint-dir> [--steps N] [--backend auto|webgpu|hip] [--case drop|dam_break] "
                 "[--out field.npy] [--precision f32|bf16]\n");
    std::exit(2);
}

Args parse(int argc, char** argv) {
    Args a;
    for (int i = 1; i < argc; i++) {
        std::string k = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) usage();
            return argv[++i];
        };
        if (k == "--model") a.model = next();
        else if (k == "--steps") a.steps = std::atoi(next().c_str());
        else if (k == "--backend") a.backend = next();
        else if (k == "--case") a.kase = next();
        else if (k == "--out") a.out = next();
        else if (k == "--precision") a.precision = next();
        else {
            std::fprintf(stderr, "unrecognized argument `%s`\n", k.c_str());
            std::exit(2);
        }
    }
    if (a.model.empty()) usage();
    return a;
}

// main.rs:63-78
void apply_case(const std::string& c, solvers::State& s, const solvers::Params& p) {
    if (c == "drop") {
        s.add_drop(p, p.nx / 2.0f, p.ny / 2.0f, 0.05f, 4.0f);
    } else if (c == "dam_break") {
        for (int y = 0; y < p.ny; y++)
            for (int x = 0; x < p.nx / 2; x++) s.h[y * p.nx + x] = p.depth + 0.15f;
    } else {
        std::fprintf(stderr, "unrecognized --case `%s` (want drop | dam_break)\n", c.c_str());
        std::exit(2);
    }
}

// main.rs:80-97: NPY v1.0, f32, 2-D, C order.
void write_npy(const std::string& path, const std::vector<float>& d, int rows, int cols) {
    std::string h = "{'descr': '<f4', 'fortran_order': False, 'shape': (" + std::to_string(rows) + ", " +
                    std::to_string(cols) + "), }";
    const size_t pad = (64 - (10 + h.size() + 1) % 64) % 64;
    h += std::string(pad, ' ') + "\n";
    std::ofstream f(path, std::ios::binary);
    f.write("\x93NUMPY\x01\x00", 8);
    const uint16_t hl = (uint16_t)h.size();
    f.write(reinterpret_cast<const char*>(&hl), 2);
    f.write(h.data(), (std::streamsize)h.size());
    f.write(reinterpret_cast<const char*>(d.data()), (std::streamsize)(d.size() * 4));
    if (!f) {
        std::fprintf(stderr, "writing .npy failed\n");
        std::exit(1);
    }
}

}  // namespace

int main(int argc, char** argv) {
    Args a = parse(argc, argv);
    if (a.backend == "cpu") {
        std::fprintf(stderr, "--backend cpu: the CPU engine is the reference's own; this binary runs the HIP engine\n");
        return 2;
    }
    if (a.backend != "auto" && a.backend != "webgpu" && a.backend != "hip") {
        std::fprintf(stderr, "unrecognized --backend `%s` (want auto | webgpu | hip; cpu is the reference's own)\n",
                     a.backend.c_str());
        return 2;
    }
    const int prec = a.precision == "bf16" ? J2K_PREC_BF16 : J2K_PREC_F32;
    try {
        engine::Model model = engine::Model::load(a.model, prec);
        const j2k_config& c = model.cfg();
        if (c.form != J2K_FORM_PATCH2D) {
            std::fprintf(stderr, "rollout runs patch2d checkpoints\n");
            return 2;
        }
        solvers::Params p;
        p.nx = c.nx;
        p.ny = c.ny;
        p.dx = c.dx;
        p.dt = c.dt;
        p.g = c.g;
        p.depth = c.depth;
        p.nu = c.nu;
        std::printf("loaded %s (%dx%d, dim %d, %d layers, mask %s) on hip\n", a.model.c_str(), c.nx, c.ny, c.dim,
                    c.layers, c.mask_kind == J2K_MASK_FULL ? "Full" : "Stencil");
        solvers::State truth = solvers::State::flat(p);
        apply_case(a.kase, truth, p);
        const auto meta = model.meta();
        const int n = p.cells();
        const engine::Normalizer norm = model.norm();
        std::vector<float> x;
        norm.encode(truth.h, truth.u, truth.v, x);
        solvers::ShallowWater solver(p);
        std::vector<float> corr((size_t)std::max(a.steps, 0));
        const auto t0 = std::chrono::steady_clock::now();
        engine::check(j2k_rollout_truth(model.handle(), solver.handle(), x.data(), meta.data(), truth.h.data(),
                                        truth.u.data(), truth.v.data(), a.steps, corr.data()));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (int step = 1; step <= a.steps; step++) std::printf("step %4d corr=%.5f\n", step, corr[step - 1]);
        std::printf("%d steps in %.1f ms (%.2f ms/step, hip)\n", a.steps, ms, ms / (a.steps > 0 ? a.steps : 1));
        if (!a.out.empty()) {
            std::vector<float> model_h(n);
            for (int i = 0; i < n; i++) model_h[i] = x[i] * norm.std[0] + norm.mean[0];  // metrics.rs decode_h
            write_npy(a.out, model_h, p.ny, p.nx);
            std::printf("wrote %s\n", a.out.c_str());
        }
    } catch (const engine::Error& e) {
        std::fprintf(stderr, "%s (status %d)\n", e.what(), e.code);
        return 1;
    }
    return 0;
}
