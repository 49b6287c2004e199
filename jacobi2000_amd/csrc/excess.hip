// The excess-energy band of training (training/jacobi2000_pt/project.py:
// 66-140) on the device, for patch2d grids x [3][ny][nx] (normalized) with
// the model's normalizer and g.
//
// The excess energy of a grid splits as a(s) s^2 + q s^3 when the anomaly
// (eta, u, v) is scaled by s (eta = (h - mean h) * std, the mass fixed):
//   a = g/2 sum eta^2 + hbar/2 sum |u|^2 + k hbar sum eta div(1 u)
//   q = 1/2 sum eta |u|^2 + k sum eta div(eta u),   k = -g/2 dt/dx,
// div(f u) the zero-wall face-flux divergence (f_a + f_b)(u_a + u_b)/4 of
// the forward-backward solver. One block per grid: the mass, then every
// per-cell value in f32 as the reference's tensor ops compute it and the
// five sums in f64 (fixed order); the projection solves
// s^2 a + s^3 q = target with the reference's start and two Newton steps
// (in f64) and rescales the grids outside the band.
#include <cmath>
#include <string>
#include <vector>

#include "abi.h"
#include "model.h"

namespace j2k {
namespace {

constexpr int ENT = 256;

__device__ double ex_block_sum(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < ENT / 64; i++) t += red[i];
    return t;
}

struct ExNorm {
    float mean[3], std[3], depth, g;
};

// Grid b: writes a + q to eout[b] (when not NULL); with PROJECT also scales
// the grid when its excess lies outside [prev*r_min, prev*r_max]
// (project.py:111-140).
template <bool PROJECT>
__global__ __launch_bounds__(ENT) void k_excess(float* xs, int nx, int ny, ExNorm nm, float dt_dx,
                                                const double* prev, float r_min, float r_max, double* eout) {
    __shared__ double red[ENT / 64];
    __shared__ float s_scale;
    const int n = nx * ny;
    float* x = xs + (size_t)blockIdx.x * 3 * n;
    // project.py:57-60 mass
    double sm = 0.0;
    for (int i = threadIdx.x; i < n; i += ENT) sm += (double)x[i];
    const float m = (float)(ex_block_sum(sm, red) / (double)n);
    // project.py:66-72
    auto eta = [&](int i) { return (x[i] - m) * nm.std[0]; };
    auto uu = [&](int i) { return x[n + i] * nm.std[1] + nm.mean[1]; };
    auto vv = [&](int i) { return x[2 * n + i] * nm.std[2] + nm.mean[2]; };
    double se2 = 0.0, ss = 0.0, ses = 0.0, so1 = 0.0, soe = 0.0;
    for (int i = threadIdx.x; i < n; i += ENT) {
        const int y = i / nx, xc = i - y * nx;
        const float e = eta(i), u = uu(i), v = vv(i);
        const float s2 = u * u + v * v;
        se2 += (double)(e * e);
        ss += (double)s2;
        ses += (double)(e * s2);
        if (dt_dx != 0.0f) {
            // project.py:75-81: the fluxes (f_a + f_b)(u_a + u_b)/4 through the
            // cell's east, west, north and south faces (zero at a wall), for
            // f = 1 and f = eta; div = ((east - west) + north) - south
            float e1 = 0.0f, ee = 0.0f, w1 = 0.0f, we = 0.0f, n1 = 0.0f, ne = 0.0f, s1 = 0.0f, se = 0.0f;
            if (xc + 1 < nx) {
                const float us = u + uu(i + 1);
                e1 = (1.0f + 1.0f) * us * 0.25f;
                ee = (e + eta(i + 1)) * us * 0.25f;
            }
            if (xc > 0) {
                const float us = uu(i - 1) + u;
                w1 = (1.0f + 1.0f) * us * 0.25f;
                we = (eta(i - 1) + e) * us * 0.25f;
            }
            if (y + 1 < ny) {
                const float vs = v + vv(i + nx);
                n1 = (1.0f + 1.0f) * vs * 0.25f;
                ne = (e + eta(i + nx)) * vs * 0.25f;
            }
            if (y > 0) {
                const float vs = vv(i - nx) + v;
                s1 = (1.0f + 1.0f) * vs * 0.25f;
                se = (eta(i - nx) + e) * vs * 0.25f;
            }
            so1 += (double)(e * (e1 - w1 + n1 - s1));
            soe += (double)(e * (ee - we + ne - se));
        }
    }
    se2 = ex_block_sum(se2, red);
    ss = ex_block_sum(ss, red);
    ses = ex_block_sum(ses, red);
    so1 = ex_block_sum(so1, red);
    soe = ex_block_sum(soe, red);
    // project.py:87-101
    const double hbar = (double)(m * nm.std[0] + (nm.mean[0] + nm.depth));
    const double g = nm.g;
    double a = se2 * (0.5 * g) + hbar * ss * 0.5;
    double q = ses * 0.5;
    if (dt_dx != 0.0f) {
        const double k = -0.5 * g * (double)dt_dx;
        a += hbar * so1 * k;
        q += soe * k;
    }
    if (eout && threadIdx.x == 0) eout[blockIdx.x] = a + q;
    if (!PROJECT) return;
    // project.py:124-134
    const double pv = prev[blockIdx.x];
    const double upper = pv * (double)r_max, lower = pv * (double)r_min;
    const double e = a + q;
    if (!(e > upper || e < lower)) return;  // uniform over the block
    if (threadIdx.x == 0) {
        const double target = fmin(fmax(e, lower * (1.0 + 1e-5)), upper * (1.0 - 1e-5));
        double s = sqrt(fmax(target / fmax(e, 1e-30), 0.0));
        for (int it = 0; it < 2; it++) {
            const double f = s * s * a + s * s * s * q - target;
            const double df = s * a * 2.0 + s * s * q * 3.0;
            s = s - f / fmax(df, 1e-30);
        }
        s_scale = (float)fmin(fmax(s, 0.0), 2.0);
    }
    __syncthreads();
    const float s = s_scale;
    // project.py:135-140
    for (int i = threadIdx.x; i < n; i += ENT) {
        x[i] = (x[i] - m) * s + m;
        x[n + i] = ((x[n + i] * nm.std[1] + nm.mean[1]) * s - nm.mean[1]) / nm.std[1];
        x[2 * n + i] = ((x[2 * n + i] * nm.std[2] + nm.mean[2]) * s - nm.mean[2]) / nm.std[2];
    }
}

ExNorm ex_norm(const j2k_config& c) {
    ExNorm nm;
    for (int i = 0; i < 3; i++) nm.mean[i] = c.norm_mean[i], nm.std[i] = c.norm_std[i];
    nm.depth = c.norm_depth;
    nm.g = c.g;
    return nm;
}

Model& ex_model(void* h, const char* fn) {
    if (!h) fail(J2K_ERR_INVALID, "model handle is NULL");
    Model& m = *static_cast<Model*>(h);
    if (m.cfg.form != J2K_FORM_PATCH2D) fail(J2K_ERR_INVALID, std::string(fn) + ": wrong model form (needs patch2d)");
    return m;
}

void ex_need(const void* p, const char* what) {
    if (!p) fail(J2K_ERR_INVALID, std::string(what) + " is NULL");
}

void ex_check(int batch, float r_min, float r_max, bool project) {
    if (batch < 0) fail(J2K_ERR_INVALID, "batch must be >= 0");
    if (project && !(r_min >= 0.0f && r_max > 0.0f && r_min <= r_max))
        fail(J2K_ERR_INVALID, "excess band needs 0 <= r_min <= r_max and r_max > 0");
}

void excess_launch(Model& m, hipStream_t s, int batch, float* x, float dt_dx, const double* prev, float r_min,
                   float r_max, double* eout, bool project) {
    if (batch == 0) return;
    const j2k_config& c = m.cfg;
    if (project)
        hipLaunchKernelGGL(k_excess<true>, dim3(batch), dim3(ENT), 0, s, x, c.nx, c.ny, ex_norm(c), dt_dx, prev,
                           r_min, r_max, eout);
    else
        hipLaunchKernelGGL(k_excess<false>, dim3(batch), dim3(ENT), 0, s, x, c.nx, c.ny, ex_norm(c), dt_dx, prev,
                           r_min, r_max, eout);
    LAUNCHCHK("k_excess");
}

hipStream_t ex_stream(Model& m, void* st) { return st ? static_cast<hipStream_t>(st) : m.stream; }

}  // namespace
}  // namespace j2k

using namespace j2k;

extern "C" {

int j2k_excess_energy_dev(void* h, int batch, const float* d_x, float dt_dx, double* d_out, void* st) {
    return guard([&] {
        Model& m = ex_model(h, "j2k_excess_energy_dev");
        ex_check(batch, 0, 0, false);
        if (batch == 0) return;
        ex_need(d_x, "x");
        ex_need(d_out, "out");
        DeviceGuard dg(m.device);
        excess_launch(m, ex_stream(m, st), batch, const_cast<float*>(d_x), dt_dx, nullptr, 0, 0, d_out, false);
    });
}

int j2k_excess_energy(void* h, int batch, const float* x, float dt_dx, double* out) {
    return guard([&] {
        Model& m = ex_model(h, "j2k_excess_energy");
        ex_check(batch, 0, 0, false);
        if (batch == 0) return;
        ex_need(x, "x");
        ex_need(out, "out");
        DeviceGuard dg(m.device);
        hipStream_t s = m.stream;
        const size_t cells = (size_t)batch * 3 * m.cfg.nx * m.cfg.ny;
        float* dx = m.get<float>(S_X0, cells);
        double* de = m.get<double>(S_TARG, (size_t)batch);
        HIPCHK(hipMemcpyAsync(dx, x, cells * 4, hipMemcpyHostToDevice, s));
        excess_launch(m, s, batch, dx, dt_dx, nullptr, 0, 0, de, false);
        HIPCHK(hipMemcpyAsync(out, de, (size_t)batch * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

int j2k_project_excess_energy_dev(void* h, int batch, float* d_x, const double* d_prev, float dt_dx, float r_min,
                                  float r_max, void* st) {
    return guard([&] {
        Model& m = ex_model(h, "j2k_project_excess_energy_dev");
        ex_check(batch, r_min, r_max, true);
        if (batch == 0) return;
        ex_need(d_x, "x");
        ex_need(d_prev, "prev");
        DeviceGuard dg(m.device);
        excess_launch(m, ex_stream(m, st), batch, d_x, dt_dx, d_prev, r_min, r_max, nullptr, true);
    });
}

int j2k_project_excess_energy(void* h, int batch, float* x, const double* prev, float dt_dx, float r_min,
                              float r_max) {
    return guard([&] {
        Model& m = ex_model(h, "j2k_project_excess_energy");
        ex_check(batch, r_min, r_max, true);
        if (batch == 0) return;
        ex_need(x, "x");
        ex_need(prev, "prev");
        DeviceGuard dg(m.device);
        hipStream_t s = m.stream;
        const size_t cells = (size_t)batch * 3 * m.cfg.nx * m.cfg.ny;
        float* dx = m.get<float>(S_X0, cells);
        double* dp = m.get<double>(S_TARG, (size_t)batch);
        HIPCHK(hipMemcpyAsync(dx, x, cells * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dp, prev, (size_t)batch * 8, hipMemcpyHostToDevice, s));
        excess_launch(m, s, batch, dx, dt_dx, dp, r_min, r_max, nullptr, true);
        HIPCHK(hipMemcpyAsync(x, dx, cells * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

}  // extern "C"
