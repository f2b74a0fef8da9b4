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

// Unit for this workgroup's iteration `iter`, or -1 when done.  XCD x owns
// units [xr[x], xr[x+1]), ranges the host balanced by work, and its
// workgroups walk them round-robin in order, so neighbouring units (which
// share x entries) run at the same time in that XCD's L2.  No atomics: a
// shared counter serialises at memory once hundreds of workgroups pop it.
__device__ inline int seg_unit(const int32_t *xr, int iter) {
  const int x = blockIdx.x % kXcd;
  const int per = gridDim.x / kXcd;
  const int u = xr[x] + (int)(blockIdx.x / kXcd) + iter * per;
  return u < xr[x + 1] ? u : -1;
}

}  // namespace spx
