// spmm.hip -- multi-RHS SpMM  Y = alpha*A*X + beta*Y  on the segmented engine's layout.
//
// X (n-by-k) and Y (m-by-k) are column-major with leading dimensions ldx, ldy.
// A pass covers up to kSpmmPass columns and reads A once:
//  - k_seg_cols walks the units of the single-column engine (seg.hip).  A
//    unit is taken kColsChunks chunks (of kSegBlock*8 nonzeros) at a time: the
//    chunks' values, column offsets and start bits are loaded into registers
//    once, and their segment structure (which thread's pieces close, how many
//    segments each thread begins, every piece's output) is worked out once;
//  - then, column after column, the unit stages that column's window of x in
//    LDS (global_load_lds, as the single-column engine does), each thread
//    multiplies its register-held values by the gathered entries, and the
//    chunk's segmented scan and coalesced store run as in seg.hip.  Only the
//    value half of the scan (DPP steps plus one exchange of wave totals
//    through LDS) is repeated per column;
//  - a segment still open at the end of a sub-block carries into the next
//    through LDS, one slot per column.
// Long rows: slice o of column j sums into partial[j * pstride + o], folded
// per column by k_long_reduce_cols.  Every addition happens in the same
// order as in the single-column engine, so each column of Y is bitwise the
// result of the single-column execute on that column of X.
#include "seg_device.h"

namespace spx {

constexpr int kColsChunks = 2;  // chunks of a unit held in registers at a time

// Inclusive sum of cn across the wave (DPP).
template <int CTRL, int ROWS, int BANKS>
__device__ inline void cnt_scan_step(int &cn) {
  cn += __builtin_amdgcn_update_dpp(0, cn, CTRL, ROWS, BANKS, false);
}

// A lane's packed segment structure for one chunk (st): bits 0-7 the start
// bits of its 8 elements, 8-15 which of them close a segment, 16-19 its first
// inner start bit (or its element count), bit 20 a sum runs on past the lane,
// bit 21 some segment closes in the lane, bit 22 the lane holds elements.
__device__ inline uint32_t st_pack(uint32_t fb, uint32_t ends, int first, int open, int hc, int has) {
  return fb | ends << 8 | (uint32_t)first << 16 | (uint32_t)open << 20 | (uint32_t)hc << 21 | (uint32_t)has << 22;
}

// Stage x[cmin, cmin + cwin) of one column of X in LDS (as seg.hip stages
// its window): s_x[lead + k] holds window entry k.  Issued by all waves; the
// caller waits for vmcnt(0) and synchronises before reading.
__device__ inline int stage_col(const double *xj, const SegUnit &U, double *s_x, int tid, int lane, int wid) {
  const double *src = xj + U.cmin;
  const int lead = (int)(((uintptr_t)src >> 3) & 1);
  const int np = (U.cwin - lead) >> 1;
  for (int q0 = wid * kWave; q0 < np; q0 += kSegBlock) {
    const int q = q0 + lane;
    if (q < np)
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + lead + 2 * q),
                                       (__attribute__((address_space(3))) void *)(s_x + 2 * lead + 2 * q0), 16, 0, 0);
  }
  if (tid == 0 && lead && U.cwin > 0) s_x[lead] = src[0];
  if (tid == kWave && ((U.cwin - lead) & 1)) s_x[lead + U.cwin - 1] = src[U.cwin - 1];
  return lead;
}

// MC chunks of a unit in registers at a time.
template <int MC>
__global__ __launch_bounds__(kSegBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_seg_cols(SpmmArgs a) {
#pragma clang fp contract(off)  // rounds as k_seg does (run_sums, axpby)
  constexpr int LE = 8;
  constexpr int CH = kSegBlock * LE;
  constexpr int NW = kSegBlock / kWave;
  __shared__ __attribute__((aligned(16))) double s_x[kSegStage + 2];
  __shared__ double s_wv[2][NW];      // per wave: scan value (double-buffered per chunk pass)
  __shared__ int s_ws[2][MC][NW];     // per wave and chunk: closed flag (double-buffered per sub-block)
  __shared__ int s_wc[2][MC][NW];     // per wave and chunk: start-bit count
  __shared__ double s_carry[kSpmmPass];  // per column: the sum open at the end of the last sub-block
  const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
  int par = 0, sbp = 0;
  SegWalk w = seg_walk(a.xr);
  int un = seg_unit_prefetch(a.units, w.u, w.end, lane);
  for (; w.u < w.end; w.u += w.step) {
    const SegUnit U = seg_unit_take(un);
    un = seg_unit_prefetch(a.units, w.u + w.step, w.end, lane);
    const bool stg = (U.flags & SEG_STAGE) != 0;
    const bool part = (U.flags & SEG_PARTIAL) != 0;
    const int64_t ostride = part ? a.pstride : a.ldy;
    int base = U.o0;  // outputs of the segments begun before the current chunk
    for (int c0 = U.e0; c0 < U.e1; c0 += MC * CH, sbp ^= 1) {
      const int nch = (U.e1 - c0 + CH - 1) / CH < MC ? (U.e1 - c0 + CH - 1) / CH : MC;  // chunks here
      const bool more = c0 + MC * CH < U.e1;  // another sub-block of this unit follows
      // ---- the sub-block's values, window offsets and segment structure, once ----
      double vv[MC][LE];
      uint32_t cw[MC][3];  // staged units: the packed window offsets
      uint32_t st[MC];
      int excl[MC];  // output of the first segment the lane begins
#pragma unroll
      for (int q = 0; q < MC; q++) {
        const int c = c0 + q * CH, ea = c + tid * LE;
        const int nv = U.e1 - ea <= 0 ? 0 : (U.e1 - ea > LE ? LE : U.e1 - ea);
        int tcl = 1;  // the lane's last piece ends with it
        uint32_t fb = 0;
        cw[q][0] = cw[q][1] = cw[q][2] = 0;
        if (nv > 0) {  // groups of 8 are padded: a lane's 8 slots are always readable
          const int B = c + wid * kSegWaveBlock;
          const int L = seg_lanes(B, U.e1);
          Vec8<double>::load(a.val, B, L, lane, vv[q]);
          const uint32_t sb = __builtin_nontemporal_load(&a.start[ea >> 3]);
          const uint32_t sn = __builtin_nontemporal_load(&a.start[(ea + LE) >> 3]);  // always padded in
          if (stg) {
            const uint32_t *pc = a.col12 + (int64_t)(ea >> 3) * 3;
            cw[q][0] = __builtin_nontemporal_load(pc);
            cw[q][1] = __builtin_nontemporal_load(pc + 1);
            cw[q][2] = __builtin_nontemporal_load(pc + 2);
          }
          fb = nv < LE ? sb & ((1u << nv) - 1) : sb;
          tcl = ea + nv < U.e1 ? (int)(sn & 1) : 1;
        } else {
#pragma unroll
          for (int i = 0; i < LE; i++) vv[q][i] = 0.0;
        }
        const uint32_t inner = fb & ~1u;
        const int hc = nv > 0 && (inner != 0 || tcl);
        st[q] = st_pack(fb, nv > 0 ? (fb >> 1) | ((uint32_t)tcl << (nv - 1)) : 0u, inner ? __builtin_ctz(inner) : nv,
                        nv > 0 && !tcl, hc, nv > 0);
        int cn = __builtin_popcount(fb);
        cnt_scan_step<0x111, 0xf, 0xf>(cn);  // row_shr:1
        cnt_scan_step<0x112, 0xf, 0xf>(cn);  // row_shr:2
        cnt_scan_step<0x114, 0xf, 0xe>(cn);  // row_shr:4
        cnt_scan_step<0x118, 0xf, 0xc>(cn);  // row_shr:8
        cnt_scan_step<0x142, 0xa, 0xf>(cn);  // row_bcast:15
        cnt_scan_step<0x143, 0xc, 0xf>(cn);  // row_bcast:31
        excl[q] = cn - __builtin_popcount(fb);  // within the wave, for now
        const int wclosed = __ballot(hc) != 0;
        if (lane == kWave - 1) {
          s_ws[sbp][q][wid] = wclosed;
          s_wc[sbp][q][wid] = cn;
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < MC; q++) {
        int cb = base, tc = base;
#pragma unroll
        for (int v = 0; v < NW; v++) {
          const int wc = s_wc[sbp][q][v];
          if (v < wid) cb += wc;
          tc += wc;
        }
        excl[q] += cb;
        base = tc;
      }
      // ---- the columns, one after another ----
      for (int j = 0; j < a.kk; j++) {
        const double *xj = a.X + (int64_t)j * a.ldx;
        int lead = 0;  // s_x[lead + k] holds window entry k
        if (stg) {
          // every thread's gathers from s_x for the previous column ended
          // before that column's last exchange barrier
          lead = stage_col(xj, U, s_x, tid, lane, wid);
          // vmcnt(0), expcnt and lgkmcnt left open (gfx9 encoding): this
          // wave's copies have landed; the barrier covers every wave's
          __builtin_amdgcn_s_waitcnt(0x0f70);
          __syncthreads();
        }
        double carry = c0 > U.e0 ? s_carry[j] : 0.0;
        double *const out = (part ? a.partial : a.Y) + (int64_t)j * ostride;
#pragma unroll
        for (int q = 0; q < MC; q++) {
          if (q >= nch) break;  // uniform
          // The registers are re-read per column through opaque copies: what
          // the compiler derives from them (unpacked offsets, lane masks) is
          // then recomputed here instead of held across the column loop.
          uint32_t t = st[q], w0 = cw[q][0], w1 = cw[q][1], w2 = cw[q][2];
          int ex = excl[q];
          asm volatile("" : "+v"(t), "+v"(w0), "+v"(w1), "+v"(w2), "+v"(ex));
          const uint32_t f = t & 0xffu, e = (t >> 8) & 0xffu;
          const int f1 = (int)((t >> 16) & 0xfu);
          // All 8 slots are multiplied, as in seg.hip: slots past the unit's
          // end hold value 0 and column (offset) 0 and feed no stored sum.
          int32_t ci[LE];
          double xg[LE], p[LE];
          if (stg) {
            unpack12(w0, w1, w2, ci);
#pragma unroll
            for (int i = 0; i < LE; i++) xg[i] = s_x[lead + ci[i]];
          } else {
            // unstaged units (wide windows, few) reload their columns per column of X
            const int c = c0 + q * CH, B = c + wid * kSegWaveBlock;
#pragma unroll
            for (int i = 0; i < LE; i++) ci[i] = 0;
            if ((t >> 22) & 1) load8i(a.col32 + U.cmin, B - U.e0, seg_lanes(B, U.e1), lane, ci);
#pragma unroll
            for (int i = 0; i < LE; i++) xg[i] = xj[ci[i]];
          }
          run_sums(vv[q], xg, f, p);  // as in seg.hip
          // the sum the lane leaves open (only a lane holding 8 elements can)
          double v[1] = {(t >> 20) & 1 ? p[LE - 1] : 0.0};
          int sg = (t >> 21) & 1;
          scan_step_v<0x111, 0xf, 0xf, 1>(v, sg);  // row_shr:1
          scan_step_v<0x112, 0xf, 0xf, 1>(v, sg);  // row_shr:2
          scan_step_v<0x114, 0xf, 0xe, 1>(v, sg);  // row_shr:4
          scan_step_v<0x118, 0xf, 0xc, 1>(v, sg);  // row_shr:8
          scan_step_v<0x142, 0xa, 0xf, 1>(v, sg);  // row_bcast:15
          scan_step_v<0x143, 0xc, 0xf, 1>(v, sg);  // row_bcast:31
          if (lane == kWave - 1) s_wv[par][wid] = v[0];
          __syncthreads();
          double prev = carry, tot = carry;
#pragma unroll
          for (int u = 0; u < NW; u++) {
            const double wv = s_wv[par][u];
            const int ws = s_ws[sbp][q][u];
            if (u < wid) prev = wv + (ws ? 0.0 : prev);
            tot = wv + (ws ? 0.0 : tot);
          }
          const double c_incl = v[0] + (sg ? 0.0 : prev);
          double carry_in = __shfl_up(c_incl, 1, kWave);
          if (lane == 0) carry_in = prev;
          // every element that ends a segment stores its running sum (plus
          // the carry from earlier lanes for piece 0) where the segment's
          // output lies; the branch on the destination is taken per unit
          auto emit = [&](auto store) {
            int k = ex - 1;  // the output of the segment holding element i
#pragma unroll
            for (int i = 0; i < LE; i++) {
              k += (f >> i) & 1;
              if ((e >> i) & 1) store(out + k, i < f1 ? p[i] + carry_in : p[i]);
            }
          };
          if (part)
            emit([](double *o, double sm) { *o = sm; });
          else if (a.beta == 0.0)
            emit([&](double *o, double sm) { *o = a.alpha * sm; });
          else
            emit([&](double *o, double sm) { *o = axpby(a.alpha, sm, a.beta, *o); });
          carry = tot;
          par ^= 1;
        }
        // read again at this column of the next sub-block, after its barriers
        if (more && tid == 0) s_carry[j] = carry;
      }
    }
  }
}

// y(row, j) for long row `row` and column j = blockIdx.y: its slices'
// partials of column j folded as k_long_reduce folds one column.
__global__ __launch_bounds__(kBlock) void k_long_reduce_cols(const double *__restrict__ partial, int64_t pstride,
                                                             int nlong, const int32_t *__restrict__ lsp,
                                                             const int32_t *__restrict__ lseg,
                                                             const int32_t *__restrict__ rows, double alpha,
                                                             double beta, double *__restrict__ Y, int64_t ldy) {
  if ((int)blockIdx.x < nlong)
    long_reduce(partial + (int64_t)blockIdx.y * pstride, lsp, lseg, rows, blockIdx.x, alpha, beta, 1,
                Y + (int64_t)blockIdx.y * ldy);
}

__global__ void k_empty_rows_multi(const int32_t *__restrict__ rows, int64_t nrows, double beta,
                                   double *__restrict__ Y, int64_t ldy, int kk) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows * kk) return;
  const int64_t i = t % nrows, j = t / nrows;
  double *yp = Y + rows[i] + j * ldy;
  *yp = beta == 0.0 ? 0.0 : beta * *yp;
}

hipError_t launch_spmm(const SpmmArgs &a, hipStream_t st) {
  if (a.kk < 1 || a.kk > kSpmmPass) return hipErrorInvalidValue;
  static const int cap = [] {
    int nb = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_seg_cols<kColsChunks>, kSegBlock, 0) != hipSuccess ||
        nb < 1)
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
    hipLaunchKernelGGL((k_seg_cols<kColsChunks>), dim3(g), dim3(kSegBlock), 0, st, a);
  }
  if (a.nlong > 0)
    hipLaunchKernelGGL(k_long_reduce_cols, dim3(a.nlong, a.kk), dim3(kBlock), 0, st, a.partial, a.pstride, a.nlong,
                       a.lsp, a.lseg, a.lrows, a.alpha, a.beta, a.Y, a.ldy);
  return hipGetLastError();
}

}  // namespace spx
