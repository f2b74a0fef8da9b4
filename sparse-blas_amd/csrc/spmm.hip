// spmm.hip -- multi-RHS SpMM  Y = alpha*A*X + beta*Y  on the segmented engine's layout.
//
// X (n-by-k) and Y (m-by-k) are column-major with leading dimensions ldx, ldy.
// One pass covers up to 8 columns (NC = the pass width rounded up to a power
// of two; columns past the pass's kk repeat its last column and are never
// stored), so A is read once per 8 columns instead of once per column.
//
// k_seg_multi walks the units of the single-column engine (seg.hip) with the
// same chunks of 2048 nonzeros, 8 per thread, loaded once into registers:
//  - a structural scan (closed flags and start-bit counts, DPP across the
//    wave, wave totals through LDS) gives every piece its output index; it is
//    the same for all columns;
//  - per column, the thread gathers x_j at its 8 columns (straight from the
//    column-major X: no transposed copy), forms its pieces and writes the
//    ones it holds whole;
//  - one segmented value scan over the pass's columns at once joins the
//    pieces that cross threads, with the chunk's carry per column in
//    registers;
//  - the sums a chunk closes are gathered in LDS per column and stored in a
//    coalesced pass (consecutive rows of one column of Y), or where they
//    close if a chunk closes more than kSegOut of them.
// Long rows: one partial per (slice, column) at partial[slice * NC + j],
// folded per column by k_long_reduce_multi.  Every addition happens in a
// fixed order, so results are bitwise reproducible; a column's sums are
// those of the single-column engine on that column.
#include "seg_device.h"

namespace spx {

template <int NC>
__device__ inline void multi_emit(const SpmmArgs &a, bool part, int j, int o, double s) {
  if (part) {
    a.partial[(int64_t)o * NC + j] = s;
  } else if (j < a.kk) {
    double *yp = a.Y + (int64_t)j * a.ldy + o;
    *yp = a.beta == 0.0 ? a.alpha * s : a.alpha * s + a.beta * *yp;
  }
}

template <int NC>
__global__ __launch_bounds__(kSegBlock) void k_seg_multi(SpmmArgs a) {
  constexpr int LE = 8;
  constexpr int CH = kSegBlock * LE;
  constexpr int NW = kSegBlock / kWave;
  __shared__ int s_ws[2][NW];  // per wave: closed flag, start-bit count (double-buffered per chunk)
  __shared__ int s_wc[2][NW];
  __shared__ double s_wv[2][NW][NC];  // per wave and column: scan value
  __shared__ double s_out[NC][kSegOut];  // sums closed in the current chunk, by column and output
  __shared__ uint8_t s_ok[kSegOut];      // which outputs are set
  const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
  for (int t = tid; t < kSegOut; t += kSegBlock) s_ok[t] = 0;
  __syncthreads();
  const double *xc[NC];
#pragma unroll
  for (int j = 0; j < NC; j++) xc[j] = a.X + (int64_t)(j < a.kk ? j : a.kk - 1) * a.ldx;
  int par = 0;
  for (int iter = 0;; iter++) {
    const int u = seg_unit(a.xr, iter);
    if (u < 0) break;
    const SegUnit U = a.units[u];
    const bool part = (U.flags & SEG_PARTIAL) != 0;
    double carry[NC];  // per column: the sum still open at the end of the previous chunk
#pragma unroll
    for (int j = 0; j < NC; j++) carry[j] = 0.0;
    int base = U.o0;  // outputs of the segments begun before the current chunk
    for (int c = U.e0; c < U.e1; c += CH, par ^= 1) {
      const int ea = c + tid * LE;
      const int nv = U.e1 - ea <= 0 ? 0 : (U.e1 - ea > LE ? LE : U.e1 - ea);
      int32_t ci[LE];
      double vv[LE];
      uint32_t fb = 0;
      int tclosed = 1;  // the thread's last piece ends with it
      if (nv > 0) {  // groups of 8 are padded: a lane's 8 slots are always readable
        const int B = c + wid * kSegWaveBlock;
        const int L = seg_lanes(B, U.e1);
        load8i(a.col, B, L, lane, ci);
        Vec8<double>::load(a.val, B, L, lane, vv);
        fb = __builtin_nontemporal_load(&a.start[ea >> 3]);
        if (nv < LE) fb &= (1u << nv) - 1;
        if (ea + nv < U.e1) tclosed = __builtin_nontemporal_load(&a.start[(ea + LE) >> 3]) & 1;
      } else {
#pragma unroll
        for (int i = 0; i < LE; i++) {
          ci[i] = 0;
          vv[i] = 0.0;
        }
      }
      const uint32_t inner = fb & ~1u;
      const int f1 = inner ? __builtin_ctz(inner) : nv;
      const int fl = inner ? 31 - __builtin_clz(inner) : 0;
      const int has_close = nv > 0 && (inner != 0 || tclosed);
      const int cnt = __builtin_popcount(fb);

      // structure: the output index of every piece
      int sg = has_close, cn = cnt;
      scan_step_sc<0x111, 0xf, 0xf>(sg, cn);  // row_shr:1
      scan_step_sc<0x112, 0xf, 0xf>(sg, cn);  // row_shr:2
      scan_step_sc<0x114, 0xf, 0xe>(sg, cn);  // row_shr:4
      scan_step_sc<0x118, 0xf, 0xc>(sg, cn);  // row_shr:8
      scan_step_sc<0x142, 0xa, 0xf>(sg, cn);  // row_bcast:15 into rows 1, 3
      scan_step_sc<0x143, 0xc, 0xf>(sg, cn);  // row_bcast:31 into rows 2, 3
      if (lane == kWave - 1) {
        s_ws[par][wid] = sg;
        s_wc[par][wid] = cn;
      }
      __syncthreads();
      int cb = base, totc = base;
#pragma unroll
      for (int w = 0; w < NW; w++) {
        const int wc = s_wc[par][w];
        if (w < wid) cb += wc;
        totc += wc;
      }
      const int k0 = cb + cn - cnt + (int)(fb & 1) - 1;  // output of the thread's piece 0
      const int kb = base - 1;  // the chunk's outputs are [kb, totc)
      const bool via_lds = totc - kb <= kSegOut;  // uniform
      auto put = [&](int j, int k, double val) {
        if (via_lds) {
          s_out[j][k - kb] = val;
          s_ok[k - kb] = 1;
        } else {
          multi_emit<NC>(a, part, j, k, val);
        }
      };

      // per column: products, piece 0, the running piece, the whole pieces
      double first[NC], v[NC];
#pragma unroll
      for (int j = 0; j < NC; j++) {
        double p[LE];
#pragma unroll
        for (int i = 0; i < LE; i++) {
          const double t = vv[i] * xc[j][i < nv ? ci[i] : 0];
          p[i] = i < nv ? t : 0.0;
        }
        double fv = 0.0, lv = 0.0;
#pragma unroll
        for (int i = 0; i < LE; i++) {
          if (i < f1) fv += p[i];
          if (i >= fl && i < nv) lv += p[i];
        }
        first[j] = fv;
        v[j] = nv > 0 && !tclosed ? lv : 0.0;
        if (nv > 0) {
          double acc = 0.0;
          int open = 0, k = k0;
#pragma unroll
          for (int i = 1; i < LE; i++) {
            if (i < nv) {
              if ((inner >> i) & 1) {
                if (open) put(j, k, acc);
                k++;
                acc = 0.0;
                open = 1;
              }
              if (open) acc += p[i];
            }
          }
          if (open && tclosed) put(j, k, acc);
        }
      }

      // values: one segmented scan of the running pieces, all columns at once
      int sv = has_close;
      scan_step_v<0x111, 0xf, 0xf>(v, sv);
      scan_step_v<0x112, 0xf, 0xf>(v, sv);
      scan_step_v<0x114, 0xf, 0xe>(v, sv);
      scan_step_v<0x118, 0xf, 0xc>(v, sv);
      scan_step_v<0x142, 0xa, 0xf>(v, sv);
      scan_step_v<0x143, 0xc, 0xf>(v, sv);
      if (lane == kWave - 1) {
#pragma unroll
        for (int j = 0; j < NC; j++) s_wv[par][wid][j] = v[j];
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < NC; j++) {
        double prev = carry[j], tot = carry[j];
#pragma unroll
        for (int w = 0; w < NW; w++) {
          const double wv = s_wv[par][w][j];
          const int ws = s_ws[par][w];
          if (w < wid) prev = wv + (ws ? 0.0 : prev);
          tot = wv + (ws ? 0.0 : tot);
        }
        const double c_incl = v[j] + (sv ? 0.0 : prev);
        double carry_in = __shfl_up(c_incl, 1, kWave);
        if (lane == 0) carry_in = prev;
        // piece 0 closes the segment open before this thread (or one that
        // begins at its first product)
        if (has_close) put(j, k0, first[j] + carry_in);
        carry[j] = tot;
      }
      if (via_lds) {
        // s_out and s_ok are written again only after the next chunk's first
        // barrier, which every thread reaches after this pass
        __syncthreads();
        for (int t = tid; t < totc - kb; t += kSegBlock) {
          if (s_ok[t]) {
#pragma unroll
            for (int j = 0; j < NC; j++) multi_emit<NC>(a, part, j, kb + t, s_out[j][t]);
            s_ok[t] = 0;
          }
        }
      }
      base = totc;
    }
  }
}

// One workgroup per long row: y(row, j) = alpha * (sum of its slices' partial
// column j, in slot order per thread, then a fixed tree) + beta * y(row, j),
// for j < kk.
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
  // up to U slices per group per round: their slot indices first, then their
  // partials, so a round costs two dependent memory round trips
  constexpr int U = 16;
  const int q1 = lsp[r + 1];
  double s = 0.0;
  for (int q = lsp[r] + g; q < q1; q += U * G) {
    int sl[U];
#pragma unroll
    for (int i = 0; i < U; i++) sl[i] = q + i * G < q1 ? lseg[q + i * G] : -1;
    double pv[U];
#pragma unroll
    for (int i = 0; i < U; i++) pv[i] = sl[i] >= 0 ? partial[(int64_t)sl[i] * K + j] : 0.0;
#pragma unroll
    for (int i = 0; i < U; i++) s += pv[i];
  }
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

template <int NC>
static hipError_t spmm_pass(const SpmmArgs &a, hipStream_t st) {
  static const int cap = [] {
    int nb = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_seg_multi<NC>, kSegBlock, 0) != hipSuccess || nb < 1)
      nb = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    const int c = nb * cus / kXcd * kXcd;
    return c < kXcd ? kXcd : c;
  }();
  if (a.nempty > 0) {
    const int64_t t = a.nempty * a.kk;
    hipLaunchKernelGGL(k_empty_rows_multi, dim3((unsigned)((t + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, a.empty,
                       a.nempty, a.beta, a.Y, a.ldy, a.kk);
  }
  if (a.nunits > 0) {
    int g = direct_grid(a.nunits);
    if (g > cap) g = cap;
    hipLaunchKernelGGL((k_seg_multi<NC>), dim3(g), dim3(kSegBlock), 0, st, a);
  }
  if (a.nlong > 0)
    hipLaunchKernelGGL((k_long_reduce_multi<NC>), dim3(a.nlong), dim3(kBlock), 0, st, a.partial, a.nlong, a.lsp,
                       a.lseg, a.lrows, a.alpha, a.beta, a.Y, a.ldy, a.kk);
  return hipGetLastError();
}

int spmm_width(int kk) {
  int K = 1;
  while (K < kk) K <<= 1;
  return K;
}

hipError_t launch_spmm(const SpmmArgs &a, hipStream_t st) {
  if (a.kk < 1 || a.kk > kSpmmPass) return hipErrorInvalidValue;
  switch (spmm_width(a.kk)) {
    case 1: return spmm_pass<1>(a, st);
    case 2: return spmm_pass<2>(a, st);
    case 4: return spmm_pass<4>(a, st);
    case 8: return spmm_pass<8>(a, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace spx
