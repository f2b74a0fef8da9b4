// build.hip -- device-side plan construction for kernel versions 1-3.
//
// The caller's CSR arrives in device memory once (one copy per array); the
// rest of the conversion runs here, so a one-shot call is bound by the
// upload, not by host passes over the nonzeros:
//  - k_check_cols: every column index inside [0, n);
//  - k_split_long: the long rows' nonzeros move to their own compact arrays,
//    the rest close up into the regular CSR;
//  - k_long_unsorted / k_long_cuts: the long rows' column order, and where
//    each row enters each column block (binary searches);
//  - k_seg_fill: the long-row slices written in the segmented engine's
//    layout (values lane-fastest per wave block, 12-bit window offsets,
//    segment-start bits) from the host's unit table.
#include "spx_device.h"
#include "spx_build.h"

namespace spx {

static inline int nb256(int64_t n) { return (int)((n + 255) / 256); }

__global__ void k_check_cols(const int32_t *__restrict__ col, int64_t nnz, int n, int *__restrict__ bad) {
  int found = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = col[k];
    found |= (c < 0) | (c >= n);
  }
  if (found) atomicOr(bad, 1);
}

hipError_t launch_check_cols(const int32_t *col, int64_t nnz, int n, int *bad, hipStream_t st) {
  if (nnz <= 0) return hipSuccess;
  int g = nb256(nnz);
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(k_check_cols, dim3(g), dim3(256), 0, st, col, nnz, n, bad);
  return hipGetLastError();
}

// Element k of the input: inside long row j when lb[j] <= k < le[j] (long
// rows sorted by row, so by position); it moves to lrp[j] + (k - lb[j]) in
// the long arrays, else to k - (nonzeros of long rows before k) in the
// regular ones.  The long rows' ranges sit in LDS (at most kMaxLong).
template <typename VI, typename VO>
__global__ __launch_bounds__(256) void k_split_long(const int32_t *__restrict__ col, const VI *__restrict__ val,
                                                    int64_t nnz, const LongRange *__restrict__ lr, int nlong,
                                                    int32_t *__restrict__ rcol, VO *__restrict__ rval,
                                                    int32_t *__restrict__ lcol, double *__restrict__ lval) {
  __shared__ LongRange s_lr[kMaxLong];
  for (int j = threadIdx.x; j < nlong; j += blockDim.x) s_lr[j] = lr[j];
  __syncthreads();
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
    // last long row starting at or before k
    int lo = 0, hi = nlong;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s_lr[mid].b <= k)
        lo = mid + 1;
      else
        hi = mid;
    }
    const int j = lo - 1;
    const int32_t c = col[k];
    const VI v = val[k];
    if (j >= 0 && k < s_lr[j].e) {
      const int64_t d = s_lr[j].dst + (k - s_lr[j].b);
      lcol[d] = c;
      lval[d] = (double)v;
    } else {
      const int64_t before = j >= 0 ? s_lr[j].dst + (s_lr[j].e - s_lr[j].b) : 0;  // long nonzeros before k
      rcol[k - before] = c;
      rval[k - before] = (VO)v;
    }
  }
}

hipError_t launch_split_long(const int32_t *col, const double *val, int64_t nnz, const LongRange *lr, int nlong,
                             int32_t *rcol, double *rval, float *rval32, int32_t *lcol, double *lval,
                             hipStream_t st) {
  if (nnz <= 0) return hipSuccess;
  int g = nb256(nnz);
  if (g > 16384) g = 16384;
  if (rval32)
    hipLaunchKernelGGL((k_split_long<double, float>), dim3(g), dim3(256), 0, st, col, val, nnz, lr, nlong, rcol,
                       rval32, lcol, lval);
  else
    hipLaunchKernelGGL((k_split_long<double, double>), dim3(g), dim3(256), 0, st, col, val, nnz, lr, nlong, rcol,
                       rval, lcol, lval);
  return hipGetLastError();
}

// flag[r] = 1 when long row r's columns decrease somewhere: one thread per
// adjacent pair over all long rows (the row found by a binary search of lrp).
__global__ void k_long_unsorted(const int32_t *__restrict__ lcol, const int32_t *__restrict__ lrp, int nlong,
                                int *__restrict__ flag) {
  const int64_t total = lrp[nlong];
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    if (lcol[k - 1] <= lcol[k]) continue;
    int lo = 0, hi = nlong;  // the row holding element k: last r with lrp[r] <= k
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (lrp[mid] <= k)
        lo = mid + 1;
      else
        hi = mid;
    }
    const int r = lo - 1;
    if (k > lrp[r]) flag[r] = 1;  // not the first element of its row
  }
}

hipError_t launch_long_unsorted(const int32_t *lcol, const int32_t *lrp, int nlong, int *flag, hipStream_t st) {
  if (nlong <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_long_unsorted, dim3(4096), dim3(256), 0, st, lcol, lrp, nlong, flag);
  return hipGetLastError();
}

// cut[r * (nblocks + 1) + b] = first element of long row r (sorted columns)
// with column >= b * width; b = nblocks gives the row's end.
__global__ void k_long_cuts(const int32_t *__restrict__ lcol, const int32_t *__restrict__ lrp, int nlong,
                            int nblocks, int width, int32_t *__restrict__ cut) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)nblocks + 1;
  if (g >= per * nlong) return;
  const int r = (int)(g / per), b = (int)(g % per);
  int lo = lrp[r], hi = lrp[r + 1];
  if (b < nblocks) {
    const int64_t key = (int64_t)b * width;
    while (lo < hi) {
      const int mid = lo + ((hi - lo) >> 1);
      if ((int64_t)lcol[mid] < key)
        lo = mid + 1;
      else
        hi = mid;
    }
  } else {
    lo = hi;
  }
  cut[g] = lo;
}

hipError_t launch_long_cuts(const int32_t *lcol, const int32_t *lrp, int nlong, int nblocks, int width, int32_t *cut,
                            hipStream_t st) {
  const int64_t total = ((int64_t)nblocks + 1) * nlong;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_long_cuts, dim3(nb256(total)), dim3(256), 0, st, lcol, lrp, nlong, nblocks, width, cut);
  return hipGetLastError();
}

// One thread per group of 8 padded elements of the segmented layout.  The
// group's unit comes from a binary search over the units' e0, each element's
// segment from a binary search over the unit's segment starts (spos);
// segment s's element e is source element ssrc[s] + (e - spos[s]).  Every
// unit stages x (column offsets from U.cmin fit 12 bits).  The 8 elements'
// values go to their lane-fastest places (seg_pos8 / seg_pos4), their
// offsets to three packed words, their start bits to one byte.
template <typename VO>
__global__ void k_seg_fill(const SegUnit *__restrict__ units, int64_t nunits, const int32_t *__restrict__ spos,
                           const int32_t *__restrict__ ssrc, const int32_t *__restrict__ scol,
                           const double *__restrict__ sval, int64_t groups, uint8_t *__restrict__ start,
                           uint32_t *__restrict__ col12, VO *__restrict__ val) {
  const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= groups) return;
  const int64_t e0 = gi * 8;
  int64_t lo = 0, hi = nunits;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)units[mid].e0 <= e0)
      lo = mid + 1;
    else
      hi = mid;
  }
  uint8_t sb = 0;
  uint32_t w[3] = {0, 0, 0};
  if (lo > 0) {
    const SegUnit U = units[lo - 1];
    const int64_t e8 = ((int64_t)U.e1 + 7) & ~(int64_t)7;
    if (e0 < e8) {
      for (int i = 0; i < 8; i++) {
        const int64_t e = e0 + i;
        uint32_t off = 0;
        double v = 0.0;
        if (e < U.e1) {
          int a = U.s0, b = U.s0 + U.ns;
          while (a < b) {
            const int mid = (a + b) >> 1;
            if ((int64_t)spos[mid] <= e)
              a = mid + 1;
            else
              b = mid;
          }
          const int s = a - 1;
          const int64_t src = (int64_t)ssrc[s] + (e - spos[s]);
          off = (uint32_t)(scol[src] - U.cmin);
          v = sval[src];
          if (e == spos[s]) sb |= (uint8_t)(1u << i);
        }
        const int bit = i * 12;
        w[bit >> 5] |= off << (bit & 31);
        if ((bit & 31) > 20) w[(bit >> 5) + 1] |= off >> (32 - (bit & 31));
        if (sizeof(VO) == 8)
          val[seg_pos8(e, U.e0, U.e1)] = (VO)v;
        else
          val[seg_pos4(e, U.e0, U.e1)] = (VO)v;
      }
    }
  }
  start[gi] = sb;
  col12[3 * gi] = w[0];
  col12[3 * gi + 1] = w[1];
  col12[3 * gi + 2] = w[2];
}

hipError_t launch_seg_fill(const SegUnit *units, int64_t nunits, const int32_t *spos, const int32_t *ssrc,
                           const int32_t *scol, const double *sval, int64_t groups, uint8_t *start, uint32_t *col12,
                           double *val, float *val32, hipStream_t st) {
  if (groups <= 0) return hipSuccess;
  if (val32)
    hipLaunchKernelGGL(k_seg_fill<float>, dim3(nb256(groups)), dim3(256), 0, st, units, nunits, spos, ssrc, scol,
                       sval, groups, start, col12, val32);
  else
    hipLaunchKernelGGL(k_seg_fill<double>, dim3(nb256(groups)), dim3(256), 0, st, units, nunits, spos, ssrc, scol,
                       sval, groups, start, col12, val);
  return hipGetLastError();
}

__global__ void k_scale_y(double beta, double *__restrict__ y, int64_t m) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) y[i] = beta == 0.0 ? 0.0 : beta * y[i];
}

hipError_t launch_scale_y(double beta, double *y, int64_t m, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale_y, dim3(nb256(m)), dim3(256), 0, st, beta, y, m);
  return hipGetLastError();
}

// Cache flush for cold-cache timing: read `n` 16-byte words (evicting L2
// and MALL without leaving dirty lines behind) and keep one value.
__global__ void k_flush_read(const uint4 *__restrict__ p, int64_t n, unsigned *__restrict__ sink) {
  unsigned acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint4 v = p[i];  // a cached load: it must displace what L2 and MALL hold
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  if (acc == 0x9e3779b9u) *sink = acc;  // keeps the loads; practically never stores
}

hipError_t launch_flush_read(const void *buf, int64_t bytes, unsigned *sink, hipStream_t st) {
  const int64_t n = bytes / 16;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_flush_read, dim3(4096), dim3(256), 0, st, static_cast<const uint4 *>(buf), n, sink);
  return hipGetLastError();
}

}  // namespace spx
