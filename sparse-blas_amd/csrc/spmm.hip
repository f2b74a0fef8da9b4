// spmm.hip -- multi-RHS SpMM  Y = alpha*A*X + beta*Y  on the segmented engine's layout.
//
// X (n-by-k) and Y (m-by-k) are column-major with leading dimensions ldx, ldy.
// Up to 32 right-hand sides are done per pass, so A is read once per 32
// columns instead of once per column:
//  - k_xpose copies the pass's columns of X into XR, row-major with K entries
//    per column index (K = the pass width rounded up to a power of two, the
//    extra entries zero), so the K values one nonzero needs are one contiguous
//    run (64 B at K = 8, 256 B at K = 32);
//  - k_spmm gives each segment (a row, or a long row's column-block slice) to a
//    group of K lanes: lane j walks the segment's nonzeros in order and
//    accumulates val * XR[col*K + j] (the value and column loads are the same
//    address for the whole group, the XR loads one contiguous run), then writes
//    y(row, j) or the slice's partial(slice, j);
//  - k_long_reduce_multi folds each long row's slice partials per column.
// Each output is one lane's sequential sum in nonzero order (long rows: slice
// sums folded in a fixed tree order), so results are bitwise reproducible.
#include "spx_device.h"

namespace spx {

template <int K>
__global__ __launch_bounds__(kBlock) void k_xpose(const double *__restrict__ X, int64_t ldx, int n, int kk,
                                                  double *__restrict__ XR) {
  constexpr int C = 64;  // columns of A (rows of X) per workgroup
  __shared__ double s[C * (K + 1)];
  const int64_t c0 = (int64_t)blockIdx.x * C;
  for (int t = threadIdx.x; t < C * K; t += kBlock) {
    const int j = t / C, c = t % C;  // consecutive threads read consecutive rows of one column
    s[c * (K + 1) + j] = (j < kk && c0 + c < n) ? X[c0 + c + (int64_t)j * ldx] : 0.0;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < C * K; t += kBlock) {
    const int c = t / K, j = t % K;  // consecutive threads write consecutive entries of XR
    if (c0 + c < n) XR[(c0 + c) * K + j] = s[c * (K + 1) + j];
  }
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_spmm(const SegUnit *__restrict__ units, const int32_t *__restrict__ xr,
                                                 const int32_t *__restrict__ sbeg, const int32_t *__restrict__ col,
                                                 const double *__restrict__ val, const double *__restrict__ XR,
                                                 double alpha, double beta, double *__restrict__ Y, int64_t ldy,
                                                 int kk, double *__restrict__ partial) {
  constexpr int G = kBlock / K;  // lane groups per workgroup
  constexpr int UN = 4;          // nonzeros per step, loads issued before the products
  const int g = threadIdx.x / K, j = threadIdx.x % K;
  for (int iter = 0;; iter++) {
    const int u = seg_unit(xr, iter);
    if (u < 0) break;
    const SegUnit U = units[u];
    for (int s = g; s < U.ns; s += G) {
      const int sg = U.s0 + s;
      const int b = sbeg[sg];
      const int e = s + 1 < U.ns ? sbeg[sg + 1] : U.e1;
      double acc = 0.0;
      int q = b;
      for (; q + UN <= e; q += UN) {
        int32_t c[UN];
        double v[UN], xv[UN];
#pragma unroll
        for (int i = 0; i < UN; i++) {
          c[i] = col[seg_pos4(q + i, U.e0, U.e1)];
          v[i] = val[seg_pos8(q + i, U.e0, U.e1)];
        }
#pragma unroll
        for (int i = 0; i < UN; i++) xv[i] = XR[(int64_t)c[i] * K + j];
#pragma unroll
        for (int i = 0; i < UN; i++) acc += v[i] * xv[i];
      }
      for (; q < e; q++) acc += val[seg_pos8(q, U.e0, U.e1)] * XR[(int64_t)col[seg_pos4(q, U.e0, U.e1)] * K + j];
      const int o = U.o0 + s;
      if (U.flags & SEG_PARTIAL) {
        partial[(int64_t)o * K + j] = acc;
      } else if (j < kk) {
        double *yp = Y + o + (int64_t)j * ldy;
        *yp = beta == 0.0 ? alpha * acc : alpha * acc + beta * *yp;
      }
    }
  }
}

// One workgroup per long row: y(row, j) = alpha * (sum of its slices' partial
// column j) + beta * y(row, j), for j < kk.
template <int K>
__global__ __launch_bounds__(kBlock) void k_long_reduce_multi(const double *__restrict__ partial, int nlong,
                                                              const int32_t *__restrict__ lsp,
                                                              const int32_t *__restrict__ lseg,
                                                              const int32_t *__restrict__ rows, double alpha,
                                                              double beta, double *__restrict__ Y, int64_t ldy,
                                                              int kk) {
  constexpr int G = kBlock / K;
  __shared__ double s_w[kBlock];
  const int r = blockIdx.x, g = threadIdx.x / K, j = threadIdx.x % K;
  if (r >= nlong) return;
  double s = 0.0;
  for (int q = lsp[r] + g; q < lsp[r + 1]; q += G) s += partial[(int64_t)lseg[q] * K + j];
  s_w[threadIdx.x] = s;
  __syncthreads();
  for (int w = G / 2; w > 0; w >>= 1) {  // fixed tree over the groups
    if (g < w) s_w[threadIdx.x] += s_w[threadIdx.x + w * K];
    __syncthreads();
  }
  if (g == 0 && j < kk) {
    double *yp = Y + rows[r] + (int64_t)j * ldy;
    *yp = beta == 0.0 ? alpha * s_w[j] : alpha * s_w[j] + beta * *yp;
  }
}

__global__ void k_empty_rows_multi(const int32_t *__restrict__ rows, int64_t nrows, double beta,
                                   double *__restrict__ Y, int64_t ldy, int kk) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows * kk) return;
  const int64_t i = t % nrows, j = t / nrows;
  double *yp = Y + rows[i] + j * ldy;
  *yp = beta == 0.0 ? 0.0 : beta * *yp;
}

template <int K>
static hipError_t spmm_pass(const SpmmArgs &a, hipStream_t st) {
  static const int cap = [] {
    int nb = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spmm<K>, kBlock, 0) != hipSuccess || nb < 1) nb = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    const int c = nb * cus / kXcd * kXcd;
    return c < kXcd ? kXcd : c;
  }();
  if (a.n > 0)
    hipLaunchKernelGGL((k_xpose<K>), dim3((unsigned)((a.n + 63) / 64)), dim3(kBlock), 0, st, a.X, a.ldx, a.n, a.kk,
                       a.XR);
  if (a.nempty > 0) {
    const int64_t t = a.nempty * a.kk;
    hipLaunchKernelGGL(k_empty_rows_multi, dim3((unsigned)((t + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, a.empty,
                       a.nempty, a.beta, a.Y, a.ldy, a.kk);
  }
  if (a.nunits > 0) {
    int g = direct_grid(a.nunits);
    if (g > cap) g = cap;
    hipLaunchKernelGGL((k_spmm<K>), dim3(g), dim3(kBlock), 0, st, a.units, a.xr, a.sbeg, a.col, a.val, a.XR, a.alpha,
                       a.beta, a.Y, a.ldy, a.kk, a.partial);
  }
  if (a.nlong > 0)
    hipLaunchKernelGGL((k_long_reduce_multi<K>), dim3(a.nlong), dim3(kBlock), 0, st, a.partial, a.nlong, a.lsp,
                       a.lseg, a.lrows, a.alpha, a.beta, a.Y, a.ldy, a.kk);
  return hipGetLastError();
}

int spmm_width(int kk) {
  int K = 1;
  while (K < kk) K <<= 1;
  return K;
}

hipError_t launch_spmm(const SpmmArgs &a, hipStream_t st) {
  switch (spmm_width(a.kk)) {
    case 1: return spmm_pass<1>(a, st);
    case 2: return spmm_pass<2>(a, st);
    case 4: return spmm_pass<4>(a, st);
    case 8: return spmm_pass<8>(a, st);
    case 16: return spmm_pass<16>(a, st);
    case 32: return spmm_pass<32>(a, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace spx
