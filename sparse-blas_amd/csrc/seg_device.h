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

// The 12-bit column offsets of the 8 elements from ea (a multiple of 8):
// three consecutive 32-bit words, unpacked to int.
__device__ inline void load8u12(const uint32_t *col12, int ea, int32_t *o) {
  const uint32_t *p = col12 + (int64_t)(ea >> 3) * 3;
  const uint32_t d0 = __builtin_nontemporal_load(p), d1 = __builtin_nontemporal_load(p + 1),
                 d2 = __builtin_nontemporal_load(p + 2);
  o[0] = (int32_t)(d0 & 0xfffu);
  o[1] = (int32_t)((d0 >> 12) & 0xfffu);
  o[2] = (int32_t)(__builtin_amdgcn_alignbit(d1, d0, 24) & 0xfffu);
  o[3] = (int32_t)((d1 >> 4) & 0xfffu);
  o[4] = (int32_t)((d1 >> 16) & 0xfffu);
  o[5] = (int32_t)(__builtin_amdgcn_alignbit(d2, d1, 28) & 0xfffu);
  o[6] = (int32_t)((d2 >> 8) & 0xfffu);
  o[7] = (int32_t)(d2 >> 20);
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

}  // namespace spx
