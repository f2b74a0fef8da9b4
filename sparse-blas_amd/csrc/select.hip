// select.hip -- the learned schedule selector's device side.
//
// Features: 32 numbers describing a CSR matrix (row-length statistics, a
// log2 row-length histogram, the share of nonzeros in long rows and how far
// columns sit from the diagonal), computed on the device from the row
// pointer and a strided sample of the columns.  Every accumulation is in
// integers (fixed point for the log distances), so the features are exact
// and independent of the reduction order.
//
// Model: a 32 -> 64 -> 64 -> 96 multilayer perceptron (ReLU, FP32) whose 96
// outputs score the schedules (kernel, tile width, partition count).  Hand
// written kernels run inference, the forward and backward passes of a
// softmax cross-entropy loss against target distributions over schedules,
// the per-parameter gradient sums and an Adam step.  Gradient sums loop over
// the batch in order (no atomics), so training is deterministic.
#include "spx_device.h"

namespace spx {

// ---- features -------------------------------------------------------------------

enum {  // per-block integer accumulators (uint64)
  FA_SUM = 0, FA_SUM2 = 1, FA_LONG = 2, FA_NB = 3,  // FA_NB..FA_NB+3: nnz by row-length class
  FA_EMPTY = 7, FA_MAX = 8, FA_HIST = 9,            // FA_HIST..+15: rows by floor(log2 len)
  FA_LOC = 25, FA_NEAR1 = 26, FA_NEAR2 = 27, FA_SAMPLES = 28, FA_N = 32
};

__device__ inline int len_class(int64_t L) { return L < 4 ? 0 : (L < 32 ? 1 : (L < 1024 ? 2 : 3)); }

__global__ __launch_bounds__(kBlock) void k_feat_partial(const int32_t *__restrict__ rowptr,
                                                         const int32_t *__restrict__ col, int m, int n,
                                                         int64_t nnz, int64_t long_th, int64_t stride,
                                                         unsigned long long *__restrict__ out) {
  __shared__ unsigned long long acc[FA_N];
  const int tid = threadIdx.x;
  if (tid < FA_N) acc[tid] = 0;
  __syncthreads();
  unsigned long long s = 0, s2 = 0, lg = 0, nb[4] = {0, 0, 0, 0}, empty = 0, mx = 0;
  const int64_t step = (int64_t)gridDim.x * kBlock;
  for (int64_t r = (int64_t)blockIdx.x * kBlock + tid; r < m; r += step) {
    const int64_t L = (int64_t)rowptr[r + 1] - rowptr[r];
    s += L;
    s2 += (unsigned long long)L * L;
    if (L >= long_th) lg += L;
    if (L == 0) {
      empty++;
    } else {
      nb[len_class(L)] += L;
      int b = 63 - __builtin_clzll((unsigned long long)L);  // floor(log2 L)
      atomicAdd(&acc[FA_HIST + (b < 15 ? b : 15)], 1ull);
    }
    if ((unsigned long long)L > mx) mx = L;
  }
  unsigned long long loc = 0, near1 = 0, near2 = 0, ns = 0;
  const int64_t nsamp = stride > 0 ? (nnz + stride - 1) / stride : 0;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + tid; t < nsamp; t += step) {
    const int64_t k = t * stride;
    int lo = 0, hi = m;  // row holding nonzero k: upper_bound(rowptr, k) - 1
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((int64_t)rowptr[mid + 1] <= k)
        lo = mid + 1;
      else
        hi = mid;
    }
    const int64_t diag = m > 0 ? (int64_t)lo * n / m : 0;
    const int64_t d = (int64_t)col[k] > diag ? (int64_t)col[k] - diag : diag - (int64_t)col[k];
    loc += (unsigned long long)(__log2f((float)d + 1.0f) * 1024.0f);
    near1 += d < 1024;
    near2 += d < 32768;
    ns++;
  }
  atomicAdd(&acc[FA_SUM], s);
  atomicAdd(&acc[FA_SUM2], s2);
  atomicAdd(&acc[FA_LONG], lg);
  for (int i = 0; i < 4; i++) atomicAdd(&acc[FA_NB + i], nb[i]);
  atomicAdd(&acc[FA_EMPTY], empty);
  atomicMax(&acc[FA_MAX], mx);
  atomicAdd(&acc[FA_LOC], loc);
  atomicAdd(&acc[FA_NEAR1], near1);
  atomicAdd(&acc[FA_NEAR2], near2);
  atomicAdd(&acc[FA_SAMPLES], ns);
  __syncthreads();
  if (tid < FA_N) out[(int64_t)blockIdx.x * FA_N + tid] = acc[tid];
}

__global__ __launch_bounds__(kBlock) void k_feat_final(const unsigned long long *__restrict__ part, int nblocks,
                                                       int m, int n, int64_t nnz, float *__restrict__ feat) {
  __shared__ unsigned long long tot[FA_N];
  const int tid = threadIdx.x;
  if (tid < FA_N) {
    unsigned long long t = 0;
    for (int b = 0; b < nblocks; b++) {
      const unsigned long long v = part[(int64_t)b * FA_N + tid];
      t = tid == FA_MAX ? (v > t ? v : t) : t + v;
    }
    tot[tid] = t;
  }
  __syncthreads();
  if (tid != 0) return;
  const double M = m > 0 ? m : 1, Z = nnz > 0 ? (double)nnz : 1.0;
  const double mean = (double)tot[FA_SUM] / M;
  const double var = fmax(0.0, (double)tot[FA_SUM2] / M - mean * mean);
  const double sd = sqrt(var);
  const double S = tot[FA_SAMPLES] > 0 ? (double)tot[FA_SAMPLES] : 1.0;
  float f[32];
  f[0] = (float)(log2((double)m + 1.0) / 24.0);
  f[1] = (float)(log2((double)n + 1.0) / 24.0);
  f[2] = (float)(log2((double)nnz + 1.0) / 32.0);
  f[3] = (float)(log2(mean + 1.0) / 12.0);
  f[4] = (float)fmin(sd / fmax(mean, 1e-30), 64.0) / 8.0f;
  f[5] = (float)(log2((double)tot[FA_MAX] + 1.0) / 24.0);
  f[6] = (float)((double)tot[FA_EMPTY] / M);
  f[7] = (float)((double)tot[FA_LONG] / Z);
  for (int i = 0; i < 16; i++) f[8 + i] = (float)((double)tot[FA_HIST + i] / M);
  for (int i = 0; i < 4; i++) f[24 + i] = (float)((double)tot[FA_NB + i] / Z);
  f[28] = (float)((double)tot[FA_LOC] / 1024.0 / S / 24.0);
  f[29] = (float)((double)tot[FA_NEAR1] / S);
  f[30] = (float)((double)tot[FA_NEAR2] / S);
  f[31] = (float)(log2(sd + 1.0) / 12.0);
  for (int i = 0; i < 32; i++) feat[i] = f[i];
}

// ---- the MLP ----------------------------------------------------------------------
// Parameter layout (FP32): W1 [64][32], b1 [64], W2 [64][64], b2 [64],
// W3 [96][64], b3 [96]; W[o][i] multiplies input i into output o.
constexpr int kIn = 32, kH = 64, kOut = 96;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kOut * kH, kParams = oB3 + kOut;

// One workgroup per sample: h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2),
// z = W3 h2 + b3.  Hidden activations are kept for the backward pass.
__global__ __launch_bounds__(128) void k_mlp_fwd(const float *__restrict__ p, const float *__restrict__ X, int B,
                                                 float *__restrict__ H1, float *__restrict__ H2,
                                                 float *__restrict__ Zo) {
  __shared__ float sx[kIn], s1[kH], s2[kH];
  const int b = blockIdx.x, t = threadIdx.x;
  if (b >= B) return;
  if (t < kIn) sx[t] = X[(int64_t)b * kIn + t];
  __syncthreads();
  if (t < kH) {
    float a = p[oB1 + t];
    for (int i = 0; i < kIn; i++) a += p[oW1 + t * kIn + i] * sx[i];
    a = a > 0.0f ? a : 0.0f;
    s1[t] = a;
    if (H1) H1[(int64_t)b * kH + t] = a;
  }
  __syncthreads();
  if (t < kH) {
    float a = p[oB2 + t];
    for (int i = 0; i < kH; i++) a += p[oW2 + t * kH + i] * s1[i];
    a = a > 0.0f ? a : 0.0f;
    s2[t] = a;
    if (H2) H2[(int64_t)b * kH + t] = a;
  }
  __syncthreads();
  if (t < kOut) {
    float a = p[oB3 + t];
    for (int i = 0; i < kH; i++) a += p[oW3 + t * kH + i] * s2[i];
    Zo[(int64_t)b * kOut + t] = a;
  }
}

// One workgroup per sample: softmax of z, its cross-entropy against the
// target distribution T, and the gradients of (loss / B) with respect to z,
// h2 and h1 (through the ReLUs).
__global__ __launch_bounds__(128) void k_mlp_bwd(const float *__restrict__ p, const float *__restrict__ H1,
                                                 const float *__restrict__ H2, const float *__restrict__ Zi,
                                                 const float *__restrict__ T, int B, float *__restrict__ dZ,
                                                 float *__restrict__ dH2, float *__restrict__ dH1,
                                                 float *__restrict__ loss) {
  __shared__ float sz[kOut], sd[kOut], s2[kH];
  __shared__ float s_mx, s_sum;
  const int b = blockIdx.x, t = threadIdx.x;
  if (b >= B) return;
  if (t < kOut) sz[t] = Zi[(int64_t)b * kOut + t];
  __syncthreads();
  if (t == 0) {
    float mx = sz[0];
    for (int i = 1; i < kOut; i++) mx = fmaxf(mx, sz[i]);
    float s = 0.0f;
    for (int i = 0; i < kOut; i++) s += __expf(sz[i] - mx);
    s_mx = mx;
    s_sum = s;
  }
  __syncthreads();
  if (t < kOut) {
    const float pr = __expf(sz[t] - s_mx) / s_sum;
    const float g = (pr - T[(int64_t)b * kOut + t]) / (float)B;
    sd[t] = g;
    dZ[(int64_t)b * kOut + t] = g;
  }
  __syncthreads();
  if (t == 0 && loss) {  // -sum T log softmax
    float l = 0.0f;
    const float lse = s_mx + __logf(s_sum);
    for (int i = 0; i < kOut; i++) l -= T[(int64_t)b * kOut + i] * (sz[i] - lse);
    loss[b] = l;
  }
  if (t < kH) {
    float a = 0.0f;
    for (int o = 0; o < kOut; o++) a += p[oW3 + o * kH + t] * sd[o];
    a = H2[(int64_t)b * kH + t] > 0.0f ? a : 0.0f;
    s2[t] = a;
    dH2[(int64_t)b * kH + t] = a;
  }
  __syncthreads();
  if (t < kH) {
    float a = 0.0f;
    for (int o = 0; o < kH; o++) a += p[oW2 + o * kH + t] * s2[o];
    dH1[(int64_t)b * kH + t] = H1[(int64_t)b * kH + t] > 0.0f ? a : 0.0f;
  }
}

// One thread per parameter: gradient = sum over the batch, in sample order.
__global__ __launch_bounds__(kBlock) void k_mlp_grad(const float *__restrict__ X, const float *__restrict__ H1,
                                                     const float *__restrict__ H2, const float *__restrict__ dZ,
                                                     const float *__restrict__ dH2, const float *__restrict__ dH1,
                                                     int B, float *__restrict__ g) {
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= kParams) return;
  float a = 0.0f;
  if (q < oB1) {
    const int o = (q - oW1) / kIn, i = (q - oW1) % kIn;
    for (int b = 0; b < B; b++) a += dH1[(int64_t)b * kH + o] * X[(int64_t)b * kIn + i];
  } else if (q < oW2) {
    const int o = q - oB1;
    for (int b = 0; b < B; b++) a += dH1[(int64_t)b * kH + o];
  } else if (q < oB2) {
    const int o = (q - oW2) / kH, i = (q - oW2) % kH;
    for (int b = 0; b < B; b++) a += dH2[(int64_t)b * kH + o] * H1[(int64_t)b * kH + i];
  } else if (q < oW3) {
    const int o = q - oB2;
    for (int b = 0; b < B; b++) a += dH2[(int64_t)b * kH + o];
  } else if (q < oB3) {
    const int o = (q - oW3) / kH, i = (q - oW3) % kH;
    for (int b = 0; b < B; b++) a += dZ[(int64_t)b * kOut + o] * H2[(int64_t)b * kH + i];
  } else {
    const int o = q - oB3;
    for (int b = 0; b < B; b++) a += dZ[(int64_t)b * kOut + o];
  }
  g[q] = a;
}

__global__ __launch_bounds__(kBlock) void k_adam(float *__restrict__ p, float *__restrict__ m1,
                                                 float *__restrict__ m2, const float *__restrict__ g, int step,
                                                 float lr, float wd) {
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= kParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gq = g[q] + wd * p[q];
  const float a = b1 * m1[q] + (1.0f - b1) * gq;
  const float v = b2 * m2[q] + (1.0f - b2) * gq * gq;
  m1[q] = a;
  m2[q] = v;
  const float ah = a / (1.0f - powf(b1, (float)step)), vh = v / (1.0f - powf(b2, (float)step));
  p[q] -= lr * ah / (sqrtf(vh) + eps);
}

// Argmax of each row of Z among the allowed schedules (mask[o] != 0),
// lowest index on ties.
__global__ void k_argmax(const float *__restrict__ Z, int B, const int32_t *__restrict__ mask,
                         int32_t *__restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  int best = -1;
  float bv = 0.0f;
  for (int o = 0; o < kOut; o++) {
    if (mask && !mask[o]) continue;
    const float v = Z[(int64_t)b * kOut + o];
    if (best < 0 || v > bv) {
      best = o;
      bv = v;
    }
  }
  out[b] = best;
}

int mlp_param_count() { return kParams; }

hipError_t launch_features(const int32_t *rowptr, const int32_t *col, int m, int n, int64_t nnz, int64_t long_th,
                           unsigned long long *part, int nblocks, float *feat, hipStream_t st) {
  int64_t stride = nnz > (1 << 20) ? (nnz + (1 << 20) - 1) / (1 << 20) : 1;
  if (nnz == 0) stride = 0;
  hipLaunchKernelGGL(k_feat_partial, dim3(nblocks), dim3(kBlock), 0, st, rowptr, col, m, n, nnz, long_th, stride,
                     part);
  hipLaunchKernelGGL(k_feat_final, dim3(1), dim3(kBlock), 0, st, part, nblocks, m, n, nnz, feat);
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const float *p, const float *X, int B, float *H1, float *H2, float *Z,
                              hipStream_t st) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_fwd, dim3(B), dim3(128), 0, st, p, X, B, H1, H2, Z);
  return hipGetLastError();
}

hipError_t launch_mlp_backward(const float *p, const float *X, const float *H1, const float *H2, const float *Z,
                               const float *T, int B, float *dZ, float *dH2, float *dH1, float *loss, float *g,
                               hipStream_t st) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_bwd, dim3(B), dim3(128), 0, st, p, H1, H2, Z, T, B, dZ, dH2, dH1, loss);
  hipLaunchKernelGGL(k_mlp_grad, dim3((kParams + kBlock - 1) / kBlock), dim3(kBlock), 0, st, X, H1, H2, dZ, dH2,
                     dH1, B, g);
  return hipGetLastError();
}

hipError_t launch_adam(float *p, float *m1, float *m2, const float *g, int step, float lr, float wd,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_adam, dim3((kParams + kBlock - 1) / kBlock), dim3(kBlock), 0, st, p, m1, m2, g, step, lr, wd);
  return hipGetLastError();
}

hipError_t launch_argmax(const float *Z, int B, const int32_t *mask, int32_t *out, hipStream_t st) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_argmax, dim3((B + kBlock - 1) / kBlock), dim3(kBlock), 0, st, Z, B, mask, out);
  return hipGetLastError();
}

}  // namespace spx
