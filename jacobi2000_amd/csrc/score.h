// Scoring and the solver as used by the model's own loops (api.hip).
#pragma once

#include "model.h"

namespace j2k {

// metrics::correlation (crates/engine/src/metrics.rs:9-23) of a*sa + ta
// against b*sb + tb over n values, into out[0] (device); one block.
void corr_affine(hipStream_t s, const float* a, float sa, float ta, const float* b, float sb, float tb, int n,
                 float* out);
// metrics.rs:78-107 on two radial profiles of n bins.
float xcorr_lag_host(const double* model, const double* solver, int n, int max_lag);
// eval-metrics' scores of `batch` trajectories (see j2k_score_dev); synchronises s.
void score_batch(Model& m, hipStream_t s, int batch, const float* d_x, const float* d_h, const float* d_u,
                 const float* d_v, const float* centers, int max_lag, j2k_eval_score* out);

// The solver handle's trajectory of one state (h, u, v device) for `steps`
// ticks: returns its frames [3][steps+1][nx*ny] in a buffer the solver owns
// (valid until its next such call); the state advances in place.
float* sw_frames_on(void* solver, hipStream_t s, float* h, float* u, float* v, int steps);
void sw_grid_of(const void* solver, int* nx, int* ny, float* depth, int* device);

}  // namespace j2k
