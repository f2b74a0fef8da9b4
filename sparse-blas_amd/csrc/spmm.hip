// spmm.hip -- multi-RHS SpMM  Y = alpha*A*X + beta*Y  on the segmented engine's layout.
//
// X (n-by-k) and Y (m-by-k) are column-major with leading dimensions ldx, ldy.
// A pass covers NC columns (NC = the pass's column count rounded up to a
// power of two, at most kSpmmPass), so A is read once per pass:
//  - k_xpose copies the pass's columns of X into XR, row-major with NC
//    entries per column index (the extra entries zero), so the NC values one
//    nonzero needs are one contiguous run (64 B at NC = 8);
//  - k_seg_multi walks the units of the single-column engine (seg.hip).  Each
//    lane takes CPL = min(NC, 4) columns of an element group of 8 consecutive
//    nonzeros; the lanes of a wave form NC/CPL column blocks of G groups
//    (lane = block*G + group), so a chunk is 4*G groups of 8 nonzeros.  One
//    load of the group's columns and values and one 8*CPL-byte read of an XR
//    row serve CPL products;
//  - per column the lane forms running sums of its products that restart at
//    its start bits, exactly as the single-column engine does; a segmented
//    scan over the block's G groups (DPP within the block's lanes, for the
//    CPL columns at once; wave totals through LDS) joins the pieces that
//    cross groups, and the open sum of each column carries to the next chunk
//    in its lane;
//  - the sums a chunk closes are gathered in LDS (up to CH/2 + 8 outputs; a
//    chunk that closes more stores each where it closes) and stored in one
//    coalesced pass: consecutive rows of one column of Y, or consecutive
//    (slice, column) partials.
// Long rows: one partial per (slice, column) at partial[slice * NC + j],
// folded per column by k_long_reduce_multi.  Every addition happens in a
// fixed order, so results are bitwise reproducible.
#include "seg_device.h"

namespace spx {

template <int NC>
__global__ __launch_bounds__(kBlock) void k_xpose(const double *__restrict__ X, int64_t ldx, int n, int kk,
                                                  double *__restrict__ XR) {
  constexpr int C = 64;  // column indices of A (rows of X) per workgroup
  __shared__ double s[C * (NC + 1)];
  const int64_t c0 = (int64_t)blockIdx.x * C;
  for (int t = threadIdx.x; t < C * NC; t += kBlock) {
    const int j = t / C, c = t % C;  // consecutive threads read consecutive rows of one column
    s[c * (NC + 1) + j] = (j < kk && c0 + c < n) ? X[c0 + c + (int64_t)j * ldx] : 0.0;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < C * NC; t += kBlock) {
    const int c = t / NC, j = t % NC;  // consecutive threads write consecutive entries of XR
    if (c0 + c < n) XR[(c0 + c) * NC + j] = s[c * (NC + 1) + j];
  }
}

template <int NC>
__device__ inline void multi_emit(const SpmmArgs &a, bool part, int j, int o, double s) {
  if (part) {
    a.partial[(int64_t)o * NC + j] = s;
  } else if (j < a.kk) {
    double *yp = a.Y + (int64_t)j * a.ldy + o;
    *yp = a.beta == 0.0 ? a.alpha * s : a.alpha * s + a.beta * *yp;
  }
}

// One step of the per-column value scan: DPP row_shr / row_bcast within the
// wave; a lane that is closed (really, or as its column's first group) takes
// nothing from below, so no sum crosses from one column's lanes to the next.
template <int CTRL, int ROWS, int BANKS>
__device__ inline void col_scan_step(double &v, int &sg) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, BANKS, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, BANKS, false);
  const int su = __builtin_amdgcn_update_dpp(0, sg, CTRL, ROWS, BANKS, false);
  if (!sg) v += __hiloint2double(hi, lo);
  sg |= su;
}

// Inclusive count scan across the wave (DPP), valid for lanes 0..G-1.
template <int CTRL, int ROWS, int BANKS>
__device__ inline void cnt_scan_step(int &cn) {
  cn += __builtin_amdgcn_update_dpp(0, cn, CTRL, ROWS, BANKS, false);
}

template <int NC>
__global__ __launch_bounds__(kSegBlock) void k_seg_multi(SpmmArgs a) {
  constexpr int LE = 8;                  // nonzeros per group per chunk
  constexpr int CPL = NC < 4 ? NC : 4;   // columns per lane
  constexpr int LPG = NC / CPL;          // lanes per element group
  constexpr int G = kWave / LPG;         // element groups per wave
  constexpr int NW = kSegBlock / kWave;
  constexpr int CH = NW * G * LE;        // nonzeros per chunk
  constexpr int NOUT = CH / 2 + 8;       // outputs gathered in LDS (more: stored where they close)
  __shared__ int s_ws[2][NW];            // per wave: closed flag, start-bit count (double-buffered)
  __shared__ int s_wc[2][NW];
  __shared__ double s_wv[2][NW][NC];     // per wave and column: scan value
  __shared__ double s_out[NC][NOUT];     // the chunk's closed sums, by column and output
  __shared__ int s_tag[NOUT];            // s_out[.][t] is this chunk's when s_tag[t] == tag
  const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
  const int jl = lane / G, g = lane % G;  // this lane's column block and element group
  const int j0 = jl * CPL;                // its first column
  for (int t = tid; t < NOUT; t += kSegBlock) s_tag[t] = 0;
  __syncthreads();
  const double *xr = a.XR + j0;
  const uint64_t below = ~0ull >> (63 - g);  // lanes g' <= g of a column block, shifted to bit 0
  int par = 0, tag = 0;
  SegWalk w = seg_walk(a.xr);
  int un = seg_unit_prefetch(a.units, w.u, w.end, lane);  // the next unit's descriptor
  for (; w.u < w.end; w.u += w.step) {
    const SegUnit U = seg_unit_take(un);
    un = seg_unit_prefetch(a.units, w.u + w.step, w.end, lane);
    const bool part = (U.flags & SEG_PARTIAL) != 0;
    double carry[CPL];  // per column: the sum still open at the end of the previous chunk
#pragma unroll
    for (int q = 0; q < CPL; q++) carry[q] = 0.0;
    int base = U.o0;  // outputs of the segments begun before the current chunk
    for (int c = U.e0; c < U.e1; c += CH, par ^= 1) {
      tag++;
      const int ea = c + (wid * G + g) * LE;  // the group's first element
      const int nv = U.e1 - ea <= 0 ? 0 : (U.e1 - ea > LE ? LE : U.e1 - ea);
      double p[CPL][LE];
      uint32_t fb = 0;
      int tclosed = 1;  // the group's last piece ends with it
      {
        int32_t ci[LE];
        double vv[LE];
        if (nv > 0) {  // groups of 8 are padded: the 8 slots are always readable
          const int off = ea - U.e0;
          const int B = U.e0 + (off & ~(kSegWaveBlock - 1));  // its wave block and slot
          const int vl = (off & (kSegWaveBlock - 1)) >> 3;
          const int L = seg_lanes(B, U.e1);
          // values and start bits first: unpacking the columns waits for
          // their load, with the other loads of the group already in flight;
          // issued at raised wave priority, as in k_seg
          __builtin_amdgcn_s_setprio(2);
          Vec8<double>::load(a.val, B, L, vl, vv);
          const uint32_t sb = __builtin_nontemporal_load(&a.start[ea >> 3]);
          const uint32_t sn = __builtin_nontemporal_load(&a.start[(ea + LE) >> 3]);  // always padded in
          if (U.flags & SEG_STAGE) {  // 12-bit offsets from cmin
            load8u12(a.col12, ea, ci);
#pragma unroll
            for (int i = 0; i < LE; i++) ci[i] += U.cmin;
          } else {
            load8i(a.col32 + U.cmin, B - U.e0, L, vl, ci);
          }
          __builtin_amdgcn_s_setprio(0);
          fb = nv < LE ? sb & ((1u << nv) - 1) : sb;
          tclosed = ea + nv < U.e1 ? (int)(sn & 1) : 1;
        } else {
#pragma unroll
          for (int i = 0; i < LE; i++) {
            ci[i] = 0;
            vv[i] = 0.0;
          }
        }
#pragma unroll
        for (int i = 0; i < LE; i++) {  // the CPL entries of one XR row are contiguous
          const double *row = xr + (int64_t)(i < nv ? ci[i] : 0) * NC;
          double xq[CPL];
          if constexpr (CPL == 4) {
            const vd2 h0 = *reinterpret_cast<const vd2 *>(row), h1 = *reinterpret_cast<const vd2 *>(row + 2);
            xq[0] = h0.x; xq[1] = h0.y; xq[2] = h1.x; xq[3] = h1.y;
          } else if constexpr (CPL == 2) {
            const vd2 h0 = *reinterpret_cast<const vd2 *>(row);
            xq[0] = h0.x; xq[1] = h0.y;
          } else {
            xq[0] = row[0];
          }
#pragma unroll
          for (int q = 0; q < CPL; q++) {
            const double t = vv[i] * xq[q];
            p[q][i] = i < nv ? t : 0.0;
          }
        }
      }
      // running sums that restart at every start bit (as in the one-column engine)
      const uint32_t inner = fb & ~1u;
      const int f1 = inner ? __builtin_ctz(inner) : nv;
      double v[CPL];
#pragma unroll
      for (int q = 0; q < CPL; q++) {
#pragma unroll
        for (int i = 1; i < LE; i++)
          if (!((fb >> i) & 1)) p[q][i] = p[q][i - 1] + p[q][i];
        double last = p[q][0];
#pragma unroll
        for (int i = 1; i < LE; i++)
          if (i < nv) last = p[q][i];
        v[q] = nv > 0 && !tclosed ? last : 0.0;
      }
      const int has_close = nv > 0 && (inner != 0 || tclosed);
      const int cnt = __builtin_popcount(fb);
      // closed flags of this column block's groups 0..g (every block has the same)
      const uint64_t closed = __ballot(has_close);
      const int sg_real = ((closed >> (jl * G)) & below) != 0;
      // values: segmented inclusive scan over the block's groups, all CPL columns
      int sg = has_close | (g == 0);
      scan_step_v<0x111, 0xf, 0xf, CPL>(v, sg);                        // row_shr:1
      if constexpr (G > 2) scan_step_v<0x112, 0xf, 0xf, CPL>(v, sg);   // row_shr:2
      if constexpr (G > 4) scan_step_v<0x114, 0xf, 0xe, CPL>(v, sg);   // row_shr:4
      if constexpr (G > 8) scan_step_v<0x118, 0xf, 0xc, CPL>(v, sg);   // row_shr:8
      if constexpr (G > 16) scan_step_v<0x142, 0xa, 0xf, CPL>(v, sg);  // row_bcast:15
      if constexpr (G > 32) scan_step_v<0x143, 0xc, 0xf, CPL>(v, sg);  // row_bcast:31
      // start-bit counts: inclusive scan over block 0's lanes, read per group
      int cn = cnt;
      cnt_scan_step<0x111, 0xf, 0xf>(cn);
      if constexpr (G > 2) cnt_scan_step<0x112, 0xf, 0xf>(cn);
      if constexpr (G > 4) cnt_scan_step<0x114, 0xf, 0xe>(cn);
      if constexpr (G > 8) cnt_scan_step<0x118, 0xf, 0xc>(cn);
      if constexpr (G > 16) cnt_scan_step<0x142, 0xa, 0xf>(cn);
      if constexpr (G > 32) cnt_scan_step<0x143, 0xc, 0xf>(cn);
      if constexpr (LPG > 1) cn = __shfl(cn, g, kWave);
      if (g == G - 1) {
#pragma unroll
        for (int q = 0; q < CPL; q++) s_wv[par][wid][j0 + q] = v[q];
      }
      if (lane == G - 1) {
        s_ws[par][wid] = (closed & (~0ull >> (64 - G))) != 0;
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
      const int kb = base - 1;  // the chunk's outputs are [kb, totc)
      const bool via_lds = totc - kb <= NOUT;  // uniform
      const uint32_t ends = nv > 0 ? ((fb >> 1) | ((uint32_t)tclosed << (nv - 1))) & ((1u << nv) - 1) : 0u;
#pragma unroll
      for (int q = 0; q < CPL; q++) {
        double prev = carry[q], tot = carry[q];
#pragma unroll
        for (int w = 0; w < NW; w++) {
          const double wv = s_wv[par][w][j0 + q];
          const int ws = s_ws[par][w];
          if (w < wid) prev = wv + (ws ? 0.0 : prev);
          tot = wv + (ws ? 0.0 : tot);
        }
        const double c_incl = v[q] + (sg_real ? 0.0 : prev);
        double carry_in = __shfl_up(c_incl, 1, kWave);
        if (g == 0) carry_in = prev;
        carry[q] = tot;
        // every element that ends a segment emits its running sum (plus the
        // carry from earlier groups for piece 0)
        int k = cb + cn - cnt - 1;  // the output of the segment holding element i
#pragma unroll
        for (int i = 0; i < LE; i++) {
          k += (fb >> i) & 1;
          if ((ends >> i) & 1) {
            const double s = i < f1 ? p[q][i] + carry_in : p[q][i];
            if (via_lds) {
              s_out[j0 + q][k - kb] = s;
              s_tag[k - kb] = tag;
            } else {
              multi_emit<NC>(a, part, j0 + q, k, s);
            }
          }
        }
      }
      if (via_lds) {
        // s_out is written again only after the next chunk's first barrier,
        // which every thread reaches after this pass
        __syncthreads();
        const int nout = totc - kb;
        if (part) {  // partial[(kb + t) * NC + jj]: consecutive threads, consecutive slots
          for (int q = tid; q < nout * NC; q += kSegBlock) {
            const int t = q / NC, jj = q % NC;
            if (s_tag[t] == tag) multi_emit<NC>(a, true, jj, kb + t, s_out[jj][t]);
          }
        } else {  // column jj of Y: kSegBlock / NC threads per column, consecutive rows
          constexpr int TPC = kSegBlock / NC;
          const int jj = tid / TPC;
          for (int t = tid % TPC; t < nout; t += TPC)
            if (s_tag[t] == tag) multi_emit<NC>(a, false, jj, kb + t, s_out[jj][t]);
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
  if (a.n > 0)
    hipLaunchKernelGGL((k_xpose<NC>), dim3((unsigned)((a.n + 63) / 64)), dim3(kBlock), 0, st, a.X, a.ldx, a.n, a.kk,
                       a.XR);
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
    case 16: return spmm_pass<16>(a, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace spx
