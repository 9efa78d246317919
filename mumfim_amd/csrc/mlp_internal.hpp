// Internal interface of the learned surrogate (ensemble of BF16-MFMA MLPs)
// used by the C ABI sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mumfim_amd.h"

namespace mf {

// Packed parameter layout of one member (W row-major out x in).
constexpr int kIn = MF_MLP_IN, kHid = MF_MLP_HIDDEN, kOut = MF_MLP_OUT;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kIn;
constexpr int kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid * kHid;
constexpr int kOffW4 = kOffB3 + kHid, kOffB4 = kOffW4 + kOut * kHid;
constexpr int kNumParams = kOffB4 + kOut;  // 135942

// Normalisation constants as the kernels read them (FP32, in HBM).
struct MlpNorm {
  float in_mean[kIn], in_inv[kIn];
  float out_mean[8], out_std[8], out_inv[8];
};

// Enqueue ensemble (member < 0) or single-member inference on stream s.
int mlp_predict(mf_mlp *m, int member, const float *X, int64_t n, float *Y,
                float *J, float *spread, hipStream_t s);
const mf_estimator &mlp_estimator(const mf_mlp *m);
int mlp_members(const mf_mlp *m);

// Surrogate-path kernels (mlp.hip).  desc: 4 floats per RVE (node count,
// fibers per volume, mean fiber length, nonlinear fraction).
hipError_t launch_surrogate_inputs(const double *F, int nF, const float *desc,
                                   int n, float *X, hipStream_t s);
// flag[r] = 1 (spread), 2 (out of range) or 0; counts[0..1] += flags.
hipError_t launch_surrogate_flags(const float *X, const float *Y,
                                  const float *spread, int n,
                                  const mf_estimator &est, int32_t *flag,
                                  hipStream_t s);
// sigma (double) from Y, and C (double, mf_batch_tangent layout) from the
// strain Jacobian J mapped to the perturbations F_i(h) F.
hipError_t launch_surrogate_outputs(const double *F, int nF, const float *Y,
                                    const float *J, int n, double *sigma,
                                    double *C, hipStream_t s);

}  // namespace mf
