// spx_build.h -- launchers of the device-side plan construction (build.hip).
#pragma once

#include "spx_internal.h"

namespace spx {

// Long row j of the input: nonzeros [b, e), moved to the long arrays from dst.
struct LongRange {
  int64_t b, e, dst;
};

// *bad |= 1 when some column index lies outside [0, n) (bad must start at 0).
hipError_t launch_check_cols(const int32_t *col, int64_t nnz, int n, int *bad, hipStream_t st);
// Split the input into the regular CSR arrays (long rows removed; values as
// double, or as float into rval32 when it is non-null) and the long rows'
// compact arrays; lr sorted by b, nlong <= kMaxLong.
hipError_t launch_split_long(const int32_t *col, const double *val, int64_t nnz, const LongRange *lr, int nlong,
                             int32_t *rcol, double *rval, float *rval32, int32_t *lcol, double *lval,
                             hipStream_t st);
// flag[r] = 1 when long row r's columns are not sorted (flag starts at 0).
hipError_t launch_long_unsorted(const int32_t *lcol, const int32_t *lrp, int nlong, int *flag, hipStream_t st);
// Where each long row enters each column block of `width` columns.
hipError_t launch_long_cuts(const int32_t *lcol, const int32_t *lrp, int nlong, int nblocks, int width, int32_t *cut,
                            hipStream_t st);
// The segmented layout of a unit table whose units all stage x: groups = the
// padded element count / 8 (every group is written, padding as zeros).
hipError_t launch_seg_fill(const SegUnit *units, int64_t nunits, const int32_t *spos, const int32_t *ssrc,
                           const int32_t *scol, const double *sval, int64_t groups, uint8_t *start, uint32_t *col12,
                           double *val, float *val32, hipStream_t st);
// y = beta * y over m rows (beta == 0 writes zeros without reading y).
hipError_t launch_scale_y(double beta, double *y, int64_t m, hipStream_t st);

// Read `bytes` of buf (a cache flush that leaves no dirty lines).
hipError_t launch_flush_read(const void *buf, int64_t bytes, unsigned *sink, hipStream_t st);

}  // namespace spx
