// spx_device.h -- device helpers shared by the gfx950 kernel files.
#pragma once

#include "spx_internal.h"

namespace spx {

// FP64 product.
__device__ inline double prod(double v, const double *x, int c) { return v * x[c]; }

// FP32 value times an FP32 (hi, lo) pair, returned as the FP64 sum of the
// exact product v*hi (split by FMA into ph + eh) and the rounded v*lo.
__device__ inline double prod(float v, const XPair *x, int c) {
  XPair p = x[c];
  float ph = v * p.hi;
  float eh = __builtin_fmaf(v, p.hi, -ph);
  float pl = __builtin_fmaf(v, p.lo, eh);
  return (double)ph + (double)pl;
}

// The dispatcher hands workgroup b to XCD b % kXcd.  Units are numbered in
// nonzero order, so giving each XCD one contiguous range of units keeps the
// x entries a unit gathers in that XCD's own L2 (neighbouring units touch
// neighbouring columns for banded rows).

// Next unit for this workgroup, uniform across the workgroup.  Workgroups
// live for the whole launch (grid = what fits resident, a multiple of kXcd):
// short-lived workgroups cannot be dispatched fast enough to keep the CUs
// full.  Static mode: the workgroups of XCD x walk x's contiguous share of
// the units in order, stride = workgroups per XCD.  Queue mode: workgroups
// pop units from a device work queue, one counter per XCD over that XCD's
// share, stealing from the other shares when their own drains.  Returns
// >= count when done.
__device__ inline int64_t next_unit(int *queue, int64_t iter, int64_t count, int *s_slot) {
  if (!queue) {
    const int64_t per = (count + kXcd - 1) / kXcd;
    const int64_t k = (int64_t)(blockIdx.x / kXcd) + iter * (int64_t)(gridDim.x / kXcd);
    if (k >= per) return count;
    const int64_t u = (int64_t)(blockIdx.x % kXcd) * per + k;
    return u < count ? u : count;
  }
  if (threadIdx.x == 0) {
    const int64_t per = (count + kXcd - 1) / kXcd;
    const int home = blockIdx.x % kXcd;
    int64_t u = count;
    for (int k = 0; k < kXcd; k++) {
      const int x = (home + k) % kXcd;
      const int64_t lim = per * (x + 1) < count ? per * (x + 1) : count;
      if ((int64_t)per * x >= lim) continue;
      const int64_t got = per * x + atomicAdd(&queue[x], 1);
      if (got < lim) {
        u = got;
        break;
      }
    }
    *s_slot = (int)(u < count ? u : count);
  }
  __syncthreads();
  const int64_t u = *s_slot;
  __syncthreads();
  return u;
}

// The segmented engine's units for this workgroup.  XCD x owns units
// [xr[x], xr[x+1]), ranges the host balanced by work, and its
// workgroups walk them round-robin in order, so neighbouring units (which
// share x entries) run at the same time in that XCD's L2.  No atomics: a
// shared counter serialises at memory once hundreds of workgroups pop it.
struct SegWalk {
  int u, end, step;  // this workgroup's units: u, u + step, ... while below end
};
__device__ inline SegWalk seg_walk(const int32_t *xr) {
  const int x = blockIdx.x % kXcd;
  return {xr[x] + (int)(blockIdx.x / kXcd), xr[x + 1], (int)(gridDim.x / kXcd)};
}

// A unit descriptor spread over a wave, one 32-bit field per lane (lane l
// holds field l % 8).  Loaded a unit ahead into one vector register, it is
// read out with readlane when the unit starts: the load has long completed
// by then, so no wait for it delays the unit.
static_assert(sizeof(SegUnit) == 8 * sizeof(int32_t), "SegUnit is 8 int32 fields");
__device__ inline int seg_unit_prefetch(const SegUnit *units, int u, int end, int lane) {
  return u < end ? reinterpret_cast<const int32_t *>(units + u)[lane & 7] : 0;
}
__device__ inline SegUnit seg_unit_take(int v) {
  SegUnit U;
  U.e0 = __builtin_amdgcn_readlane(v, 0);
  U.e1 = __builtin_amdgcn_readlane(v, 1);
  U.s0 = __builtin_amdgcn_readlane(v, 2);
  U.ns = __builtin_amdgcn_readlane(v, 3);
  U.o0 = __builtin_amdgcn_readlane(v, 4);
  U.cmin = __builtin_amdgcn_readlane(v, 5);
  U.cwin = __builtin_amdgcn_readlane(v, 6);
  U.flags = __builtin_amdgcn_readlane(v, 7);
  return U;
}

}  // namespace spx
