// select_api.cpp -- extern "C" entry points of the learned schedule selector
// (features, MLP inference and training steps; kernels in select.hip).
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/spx_capi.h"
#include "spx_internal.h"

using namespace spx;

namespace spx {
int mlp_param_count();
hipError_t launch_features(const int32_t *rowptr, const int32_t *col, int m, int n, int64_t nnz, int64_t long_th,
                           unsigned long long *part, int nblocks, float *feat, hipStream_t st);
hipError_t launch_mlp_forward(const float *p, const float *X, int B, float *H1, float *H2, float *Z, hipStream_t st);
hipError_t launch_mlp_backward(const float *p, const float *X, const float *H1, const float *H2, const float *Z,
                               const float *T, int B, float *dZ, float *dH2, float *dH1, float *loss, float *g,
                               hipStream_t st);
hipError_t launch_adam(float *p, float *m1, float *m2, const float *g, int step, float lr, float wd, hipStream_t st);
hipError_t launch_argmax(const float *Z, int B, const int32_t *mask, int32_t *out, hipStream_t st);
}  // namespace spx

namespace {
constexpr int kFeatBlocks = 1024;
constexpr int kH = 64, kOut = 96;
int set_err(hipError_t e) {
  if (e == hipSuccess) return SPX_OK;
  set_last_error(std::string("HIP: ") + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? SPX_EMEM : SPX_EDEV;
}
}  // namespace

extern "C" {

int spx_sel_param_count(void) { return mlp_param_count(); }

int64_t spx_mlp_work_floats(int B) { return (int64_t)(B > 0 ? B : 0) * (2 * kH + kOut) * 2; }

int spx_sel_features(int m, int n, int64_t nnz, const int32_t *rowptr, const int32_t *col, int64_t long_row,
                     float *feat, void *stream) {
  if (m < 0 || n < 0 || nnz < 0 || !rowptr || !feat || (nnz > 0 && !col)) {
    set_last_error("spx_sel_features: bad arguments");
    return SPX_EARG;
  }
  unsigned long long *part = nullptr;
  hipError_t e = spx::dev_malloc((void **)&part, sizeof(unsigned long long) * 32 * kFeatBlocks);
  if (e != hipSuccess) return set_err(e);
  const int64_t th = long_row > 0 ? long_row : ((int64_t)1 << 62);
  e = launch_features(rowptr, col, m, n, nnz, th, part, kFeatBlocks, feat, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  (void)hipFree(part);
  return set_err(e);
}

int spx_mlp_forward(const float *params, const float *X, int B, float *Z, float *work, void *stream) {
  if (!params || B < 0 || (B > 0 && (!X || !Z || !work))) {
    set_last_error("spx_mlp_forward: bad arguments");
    return SPX_EARG;
  }
  float *H1 = work, *H2 = work + (int64_t)B * kH;
  return set_err(launch_mlp_forward(params, X, B, H1, H2, Z, (hipStream_t)stream));
}

int spx_mlp_train_step(float *params, float *m1, float *m2, int step, float lr, float wd, const float *X,
                       const float *T, int B, float *loss, float *grad, float *work, void *stream) {
  if (!params || !m1 || !m2 || step < 1 || B <= 0 || !X || !T || !grad || !work) {
    set_last_error("spx_mlp_train_step: bad arguments");
    return SPX_EARG;
  }
  hipStream_t st = (hipStream_t)stream;
  float *H1 = work, *H2 = H1 + (int64_t)B * kH, *Z = H2 + (int64_t)B * kH;
  float *dH1 = Z + (int64_t)B * kOut, *dH2 = dH1 + (int64_t)B * kH, *dZ = dH2 + (int64_t)B * kH;
  hipError_t e = launch_mlp_forward(params, X, B, H1, H2, Z, st);
  if (e == hipSuccess) e = launch_mlp_backward(params, X, H1, H2, Z, T, B, dZ, dH2, dH1, loss, grad, st);
  if (e == hipSuccess && lr > 0.0f) e = launch_adam(params, m1, m2, grad, step, lr, wd, st);
  return set_err(e);
}

int spx_mlp_argmax(const float *Z, int B, const int32_t *mask, int32_t *out, void *stream) {
  if (B < 0 || (B > 0 && (!Z || !out))) {
    set_last_error("spx_mlp_argmax: bad arguments");
    return SPX_EARG;
  }
  return set_err(launch_argmax(Z, B, mask, out, (hipStream_t)stream));
}

}  // extern "C"
