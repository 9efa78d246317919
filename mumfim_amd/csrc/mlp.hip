// Learned surrogate: placeholder until the MFMA kernels land.
#include "mf_internal.hpp"

extern "C" {
struct mf_mlp {
  int unused;
};
int mf_mlp_create(uint64_t, mf_mlp **out) {
  if (out) *out = nullptr;
  mf::set_error("surrogate not built");
  return MF_ERR_UNSUPPORTED;
}
void mf_mlp_destroy(mf_mlp *m) { delete m; }
int mf_mlp_num_params(const mf_mlp *) { return 0; }
int mf_mlp_set_normalization(mf_mlp *, const float *, const float *, const float *,
                             const float *) { return MF_ERR_UNSUPPORTED; }
int mf_mlp_get_weights(mf_mlp *, float *) { return MF_ERR_UNSUPPORTED; }
int mf_mlp_set_weights(mf_mlp *, const float *) { return MF_ERR_UNSUPPORTED; }
int mf_mlp_forward_device(mf_mlp *, const float *, int64_t, float *, float *) {
  return MF_ERR_UNSUPPORTED;
}
int mf_mlp_train_step_device(mf_mlp *, const float *, const float *, int64_t, float,
                             float *) { return MF_ERR_UNSUPPORTED; }
int mf_mlp_gradient_device(mf_mlp *, const float *, const float *, int64_t, float *,
                           float *) { return MF_ERR_UNSUPPORTED; }
void *mf_mlp_stream(mf_mlp *) { return nullptr; }
}
