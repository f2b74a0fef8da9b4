// spx_device.h -- device helpers shared by the gfx950 kernel files.
#pragma once

#include "spx_internal.h"

namespace spx {

// FP64 product.
__device__ inline double prod(double v, double xv) { return v * xv; }

// Precision form: an FP32 value times x split into FP32 hi = fl(x) and
// lo = fl(x - hi), returned as the FP64 sum of the exact product v*hi (split
// by FMA into ph + eh) and the rounded v*lo.  x stays FP64 in memory; the
// split costs four instructions and no extra pass over x.
__device__ inline double prod(float v, double xv) {
  const float hi = (float)xv;
  const float lo = (float)(xv - (double)hi);
  const float ph = v * hi;
  const float eh = __builtin_fmaf(v, hi, -ph);
  const float pl = __builtin_fmaf(v, lo, eh);
  return (double)ph + (double)pl;
}

template <typename V>
__device__ inline double prod(V v, const double *x, int c) {
  return prod(v, x[c]);
}


// Stage x[cmin, cmin + cwin) in LDS for a workgroup of NT threads: entries go
// straight from global memory to LDS, 16 bytes per lane
// (global_load_lds_dwordx4: lane l's 16 bytes land at the wave's LDS address
// + 16 l), all of the window at once, so the copy costs one memory latency
// and no registers.  Pairs of entries start at the first 16-byte aligned
// entry; a leading and a trailing single entry go through registers.
// Returns `lead`: s_x[lead + k] holds window entry k.  The caller waits
// (stage_wait) and then synchronises the workgroup before reading s_x, and
// synchronises before restaging (earlier gathers must be done).
template <int NT>
__device__ inline int stage_window(const double *x, int cmin, int cwin, double *s_x, int tid) {
  const double *src = x + cmin;
  const int lead = (int)(((uintptr_t)src >> 3) & 1);
  const int np = (cwin - lead) >> 1;
  const int wid = tid / kWave, lane = tid % kWave;
  for (int q0 = wid * kWave; q0 < np; q0 += NT) {
    const int q = q0 + lane;
    if (q < np)
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + lead + 2 * q),
                                       (__attribute__((address_space(3))) void *)(s_x + 2 * lead + 2 * q0), 16, 0,
                                       0);
  }
  if (tid == 0 && lead && cwin > 0) s_x[lead] = src[0];
  if (tid == kWave % NT && ((cwin - lead) & 1)) s_x[lead + cwin - 1] = src[cwin - 1];
  return lead;
}
// This wave's outstanding loads, the LDS copies included, have completed
// (vmcnt(0) with expcnt and lgkmcnt left open, gfx9 encoding).
__device__ inline void stage_wait() { __builtin_amdgcn_s_waitcnt(0x0f70); }

// A unit descriptor of up to 16 int32 fields spread over a wave, one field
// per lane (lane l holds field l % 16).  Loaded a unit ahead into one vector
// register, it is read with readlane when the unit starts, so no wait for it
// delays the unit.
template <typename T>
__device__ inline int desc_prefetch(const T *units, int64_t u, int64_t end, int lane) {
  static_assert(sizeof(T) % 4 == 0 && sizeof(T) <= 64, "descriptor of at most 16 int32 fields");
  constexpr int F = (int)(sizeof(T) / 4);
  return u < end ? reinterpret_cast<const int32_t *>(units + u)[(lane % 16) < F ? lane % 16 : 0] : 0;
}
__device__ inline int desc_field(int v, int f) { return __builtin_amdgcn_readlane(v, f); }

// Static unit walk of a persistent grid: the workgroups of XCD x (b % kXcd)
// take x's contiguous share of [0, count) in order, stride = workgroups per
// XCD, so units that share x entries run together in one L2.
struct UnitWalk {
  int64_t u, end, step;
};
__device__ inline UnitWalk unit_walk(int64_t count) {
  const int64_t per = (count + kXcd - 1) / kXcd;
  const int64_t x = blockIdx.x % kXcd;
  const int64_t b = x * per, e = b + per < count ? b + per : count;
  return {b + (int64_t)(blockIdx.x / kXcd), e, (int64_t)(gridDim.x / kXcd)};
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

// Next unit for this workgroup from a per-XCD range table (the host split
// the units into kXcd contiguous ranges of about equal work): workgroup b
// walks XCD (b mod kXcd)'s range round-robin.  Returns count when done.
__device__ inline int64_t next_unit_ranged(const int32_t *xr, int64_t iter, int64_t count) {
  const int x = blockIdx.x % kXcd;
  const int64_t u = xr[x] + (int64_t)(blockIdx.x / kXcd) + iter * (int64_t)(gridDim.x / kXcd);
  return u < xr[x + 1] ? u : count;
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
