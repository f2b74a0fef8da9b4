// seg_device.h -- device helpers shared by the segmented-engine kernels
// (seg.hip: one right-hand side; spmm.hip: several).
#pragma once

#include "spx_device.h"

namespace spx {

// A lane's 8 elements of wave block B with L lanes (the layout of
// seg_pos4 / seg_pos8): 16-byte loads whose runs are contiguous across lanes.
// The matrix is read once per SpMV, so these loads are non-temporal: they do
// not displace the x windows and y lines that other units reuse from L2.
typedef double vd2 __attribute__((ext_vector_type(2)));
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef int vi4 __attribute__((ext_vector_type(4)));

template <typename V>
struct Vec8;
template <>
struct Vec8<double> {
  __device__ static inline void load(const double *p, int64_t B, int L, int lane, double *o) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const vd2 u = __builtin_nontemporal_load(reinterpret_cast<const vd2 *>(p + B + (int64_t)k * 2 * L + 2 * lane));
      o[2 * k] = u.x;
      o[2 * k + 1] = u.y;
    }
  }
};
template <>
struct Vec8<float> {
  __device__ static inline void load(const float *p, int64_t B, int L, int lane, float *o) {
    const vf4 u = __builtin_nontemporal_load(reinterpret_cast<const vf4 *>(p + B + 4 * lane));
    const vf4 w = __builtin_nontemporal_load(reinterpret_cast<const vf4 *>(p + B + (int64_t)4 * L + 4 * lane));
    o[0] = u.x; o[1] = u.y; o[2] = u.z; o[3] = u.w; o[4] = w.x; o[5] = w.y; o[6] = w.z; o[7] = w.w;
  }
};
__device__ inline void load8i(const int32_t *p, int64_t B, int L, int lane, int32_t *o) {
  const vi4 u = __builtin_nontemporal_load(reinterpret_cast<const vi4 *>(p + B + 4 * lane));
  const vi4 w = __builtin_nontemporal_load(reinterpret_cast<const vi4 *>(p + B + (int64_t)4 * L + 4 * lane));
  o[0] = u.x; o[1] = u.y; o[2] = u.z; o[3] = u.w; o[4] = w.x; o[5] = w.y; o[6] = w.z; o[7] = w.w;
}

// The 8 12-bit column offsets packed in three 32-bit words, unpacked to int.
__device__ inline void unpack12(uint32_t d0, uint32_t d1, uint32_t d2, int32_t *o) {
  o[0] = (int32_t)(d0 & 0xfffu);
  o[1] = (int32_t)((d0 >> 12) & 0xfffu);
  o[2] = (int32_t)(__builtin_amdgcn_alignbit(d1, d0, 24) & 0xfffu);
  o[3] = (int32_t)((d1 >> 4) & 0xfffu);
  o[4] = (int32_t)((d1 >> 16) & 0xfffu);
  o[5] = (int32_t)(__builtin_amdgcn_alignbit(d2, d1, 28) & 0xfffu);
  o[6] = (int32_t)((d2 >> 8) & 0xfffu);
  o[7] = (int32_t)(d2 >> 20);
}

// The 12-bit column offsets of the 8 elements from ea (a multiple of 8):
// three consecutive 32-bit words.
__device__ inline void load8u12(const uint32_t *col12, int ea, int32_t *o) {
  const uint32_t *p = col12 + (int64_t)(ea >> 3) * 3;
  unpack12(__builtin_nontemporal_load(p), __builtin_nontemporal_load(p + 1), __builtin_nontemporal_load(p + 2), o);
}

// A lane's running sums over its 8 elements, restarting at each start bit
// (bit i of f): p[i] = v[i]*xg[i] where element i starts a segment, else
// p[i-1] + v[i]*xg[i], as one fused multiply-add for FP64 values (the
// precision form adds its split product).  Written out so that every kernel
// that forms these sums rounds them alike.
__device__ inline void run_sums(const double *v, const double *xg, uint32_t f, double *p) {
  p[0] = v[0] * xg[0];
#pragma unroll
  for (int i = 1; i < 8; i++) p[i] = __builtin_fma(v[i], xg[i], (f >> i) & 1 ? 0.0 : p[i - 1]);
}
__device__ inline void run_sums(const float *v, const double *xg, uint32_t f, double *p) {
  p[0] = prod(v[0], xg[0]);
#pragma unroll
  for (int i = 1; i < 8; i++) p[i] = ((f >> i) & 1 ? 0.0 : p[i - 1]) + prod(v[i], xg[i]);
}

// The stored result: alpha*s, or alpha*s + beta*y0 rounded once after beta*y0.
__device__ inline double axpby(double alpha, double s, double beta, double y0) {
  return beta == 0.0 ? alpha * s : __builtin_fma(alpha, s, beta * y0);
}

// One step of a wave64 inclusive scan with DPP (data-parallel primitives:
// cross-lane operands read by the ALU, no LDS traffic).  Each lane combines
// the (sum, closed, count) triple of the lane CTRL names, or the identity
// (0, 0, 0) where that lane is out of range or masked off (old = 0 and
// bound_ctrl off): the segmented sum restarts after a closing lane, the
// count always adds.  The six steps row_shr 1, 2, 4, 8 then row_bcast 15
// and 31 give the full 64-lane inclusive scan.
template <int CTRL, int ROWS, int BANKS>
__device__ inline void scan_step(double &v, int &sg, int &cn) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, BANKS, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, BANKS, false);
  const int su = __builtin_amdgcn_update_dpp(0, sg, CTRL, ROWS, BANKS, false);
  const int cu = __builtin_amdgcn_update_dpp(0, cn, CTRL, ROWS, BANKS, false);
  cn += cu;
  if (!sg) {
    v += __hiloint2double(hi, lo);
    sg |= su;
  }
}

// The structural half of scan_step: closed flags and start-bit counts only.
template <int CTRL, int ROWS, int BANKS>
__device__ inline void scan_step_sc(int &sg, int &cn) {
  const int su = __builtin_amdgcn_update_dpp(0, sg, CTRL, ROWS, BANKS, false);
  const int cu = __builtin_amdgcn_update_dpp(0, cn, CTRL, ROWS, BANKS, false);
  cn += cu;
  sg |= su;
}

// The value half of scan_step for NC sums that share one segment structure:
// every column restarts where the structure closes.
template <int CTRL, int ROWS, int BANKS, int NC>
__device__ inline void scan_step_v(double (&v)[NC], int &sg) {
  const int su = __builtin_amdgcn_update_dpp(0, sg, CTRL, ROWS, BANKS, false);
#pragma unroll
  for (int j = 0; j < NC; j++) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v[j]), CTRL, ROWS, BANKS, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v[j]), CTRL, ROWS, BANKS, false);
    if (!sg) v[j] += __hiloint2double(hi, lo);
  }
  sg |= su;
}

// Long rows: fold the per-column-block partial sums.
// The segmented engine sweeps the long rows by column blocks and leaves one
// partial sum per (column block, row) slice; lseg[lsp[r] .. lsp[r+1]) lists
// row r's slices in column order.  One workgroup per long row adds them in a
// fixed tree order, then y[row] = alpha * sum + beta * y[row] (`assign`), or
// y[row] += alpha * sum when the regular pass has already written beta * y0
// for the row (it is empty there).
__device__ inline void long_reduce(const double *__restrict__ partial, const int32_t *__restrict__ lsp,
                                   const int32_t *__restrict__ lseg, const int32_t *__restrict__ rows, int r,
                                   double alpha, double beta, int assign, double *__restrict__ y) {
  // Up to U slices per thread per round: all their slot indices are loaded
  // first, then all the partials, so a row of up to U*kBlock slices costs two
  // dependent memory round trips.  A thread adds its slices in slot order.
  constexpr int U = 16;
  __shared__ double s_w[kBlock / kWave];
  const int tid = threadIdx.x;
  const int j0 = lsp[r], j1 = lsp[r + 1];
  double s = 0.0;
  for (int j = j0 + tid; j < j1; j += U * kBlock) {
    int q[U];
#pragma unroll
    for (int k = 0; k < U; k++) q[k] = j + k * kBlock < j1 ? lseg[j + k * kBlock] : -1;
    double v[U];
#pragma unroll
    for (int k = 0; k < U; k++) v[k] = q[k] >= 0 ? partial[q[k]] : 0.0;
#pragma unroll
    for (int k = 0; k < U; k++) s += v[k];
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
  if (tid % kWave == 0) s_w[tid / kWave] = s;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / kWave; w++) t += s_w[w];
    const int row = rows[r];
    if (!assign)
      y[row] = __builtin_fma(alpha, t, y[row]);
    else
      y[row] = axpby(alpha, t, beta, y[row]);
  }
}

}  // namespace spx
