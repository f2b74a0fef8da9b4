// seg.hip -- the segmented engine: row-aligned units, 16-byte loads, LDS x.
//
// A "segment" is a run of nonzeros that sums into one output: a nonempty row
// (plain SpMV) or the slice of a long row inside one column block.  Segments
// are grouped into units at segment boundaries, so no segment is ever split
// between workgroups and no carries leave a unit.  Each unit's nonzeros are
// copied to an 8-aligned region (padding <= 7 entries), so a thread owns 8
// aligned nonzeros per chunk, stored lane-fastest per wave (seg_pos4,
// seg_pos8) so its 16-byte loads are contiguous across the wave, plus one
// byte of segment-start bits.  A unit's segments write consecutive outputs,
// so a sum's destination follows from its index without a further load.  A
// unit flagged SEG_STAGE first stages its
// column window of x (at most kSegStage entries) in LDS, and every gather
// hits LDS; a unit flagged SEG_PARTIAL stores its segment sums into a
// partial array instead of updating y.
//
// Per chunk of kSegBlock*8 nonzeros, each thread splits its 8 products at its
// start bits into pieces.  Pieces bounded on both sides inside the thread are
// whole segments and are written out at once.  The piece that continues from
// earlier threads and the piece that runs on past the thread are joined by a
// block-wide segmented scan (DPP steps across each wave, then one exchange of
// the four wave totals through LDS), which also counts start bits to give
// every piece its segment index.  The sum still open at the end of a chunk carries into
// the next chunk in registers.  One barrier per chunk: the wave totals are
// double-buffered.  Every addition happens in a fixed order, so the result
// is bitwise reproducible.
#include "seg_device.h"

namespace spx {

// Partials are read once, by the folds after the kernel: stored non-temporal,
// so they do not evict the x entries that the units' gathers reuse from L2.
template <typename V, typename X>
__device__ inline void seg_emit(const SegArgs<V, X> &a, bool part, int o, double s) {
  if (part)
    __builtin_nontemporal_store(s, &a.partial[o]);
  else
    a.y[o] = axpby(a.alpha, s, a.beta, a.beta == 0.0 ? 0.0 : a.y[o]);
}

template <typename V, typename X, bool STAGE, int NOUT>
__global__ __launch_bounds__(kSegBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_seg(SegArgs<V, X> a) {
#pragma clang fp contract(off)  // the sums round as written (run_sums, axpby): the SpMM kernel matches them
  constexpr int LE = 8;
  constexpr int CH = kSegBlock * LE;
  constexpr int NW = kSegBlock / kWave;
  __shared__ __attribute__((aligned(16))) X s_x[STAGE ? kSegStage + 2 : 2];
  __shared__ double s_wv[2][NW];  // per wave: scan value, closed flag, start-bit count
  __shared__ int s_ws[2][NW];
  __shared__ int s_wc[2][NW];
  __shared__ double s_out[NOUT];  // sums closed in the current chunk, by output
  __shared__ uint8_t s_ok[NOUT];  // which of them are set
  const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
  for (int t = tid; t < NOUT; t += kSegBlock) s_ok[t] = 0;
  __syncthreads();
  int par = 0;  // double-buffer slot of the wave totals, alternating per chunk
  SegWalk w = seg_walk(a.xr);
  // Each unit's descriptor is loaded one unit ahead, so no wait for it stands
  // between one unit's last chunk and the loads of the next unit's x window.
  int un = seg_unit_prefetch(a.units, w.u, w.end, lane);
  for (; w.u < w.end; w.u += w.step) {
    const SegUnit U = seg_unit_take(un);
    un = seg_unit_prefetch(a.units, w.u + w.step, w.end, lane);
    const bool stg = STAGE && (U.flags & SEG_STAGE);
    const bool part = (U.flags & SEG_PARTIAL) != 0;
    int lead = 0;  // s_x[lead + k] holds window entry k (lead: 0 or 1)
    if (stg) {
      // The window goes straight from global memory to LDS, 16 bytes per lane
      // (global_load_lds_dwordx4: lane l's 16 bytes land at the wave's LDS
      // address + 16 l), all of it at once: no registers, one memory
      // latency.  Pairs of 8-byte entries start at the first 16-byte
      // aligned entry; a leading and a trailing single entry are copied
      // through registers.
      static_assert(sizeof(X) == 8, "x entries are FP64");
      const X *src = a.x + U.cmin;
      lead = (int)(((uintptr_t)src >> 3) & 1);
      const int np = (U.cwin - lead) >> 1;
      __syncthreads();  // the previous unit's gathers from s_x are done
      for (int q0 = wid * kWave; q0 < np; q0 += kSegBlock) {
        const int q = q0 + lane;
        if (q < np)
          __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + lead + 2 * q),
                                           (__attribute__((address_space(3))) void *)(s_x + 2 * lead + 2 * q0), 16, 0, 0);
      }
      if (tid == 0 && lead && U.cwin > 0) s_x[lead] = src[0];
      if (tid == kWave && ((U.cwin - lead) & 1)) s_x[lead + U.cwin - 1] = src[U.cwin - 1];
      // vmcnt(0), expcnt and lgkmcnt left open (gfx9 encoding): this wave's
      // copies have landed in LDS; the barrier then covers every wave's
      __builtin_amdgcn_s_waitcnt(0x0f70);
      __syncthreads();
    }
    double carry = 0.0;  // sum still open at the end of the previous chunk
    int base = U.o0;     // outputs of the segments begun before the current chunk
    for (int c = U.e0; c < U.e1; c += CH, par ^= 1) {
      const int ea = c + tid * LE;
      const int nv = U.e1 - ea <= 0 ? 0 : (U.e1 - ea > LE ? LE : U.e1 - ea);
      double p[LE];
      uint32_t fb = 0;
      int tclosed = 1;  // the thread's last piece ends with it
      {
        int32_t ci[LE];
        V vv[LE];
        if (nv > 0) {  // groups of 8 are padded: a lane's 8 slots are always readable
          // The values and start bits are loaded first and the columns last:
          // unpacking the 12-bit offsets waits for their load, and that wait
          // then finds the other loads of the chunk already in flight.  The
          // wave issues them at raised priority (s_setprio), so waves with
          // loads to send go ahead of waves in their scan and emit steps.
          const int B = c + wid * kSegWaveBlock;  // this wave's block
          const int L = seg_lanes(B, U.e1);
          __builtin_amdgcn_s_setprio(2);
          Vec8<V>::load(a.val, B, L, lane, vv);
          const uint32_t sb = __builtin_nontemporal_load(&a.start[ea >> 3]);
          const uint32_t sn = __builtin_nontemporal_load(&a.start[(ea + LE) >> 3]);  // always padded in
          if (stg)
            load8u12(a.col12, ea, ci);  // window offsets
          else
            load8i(a.col32 + U.cmin, B - U.e0, L, lane, ci);
          __builtin_amdgcn_s_setprio(0);
          fb = nv < LE ? sb & ((1u << nv) - 1) : sb;
          tclosed = ea + nv < U.e1 ? (int)(sn & 1) : 1;
        } else {
#pragma unroll
          for (int i = 0; i < LE; i++) {
            ci[i] = 0;
            vv[i] = (V)0;
          }
        }
        // All 8 slots are gathered and multiplied: a lane's slots past the
        // unit's end hold value 0 and column (offset) 0, always in range, and
        // feed only running sums that are never stored (that lane's last
        // segment ends at the unit's end, so the lane carries nothing on).
        double xg[LE];
        if (stg) {
#pragma unroll
          for (int i = 0; i < LE; i++) xg[i] = s_x[lead + ci[i]];
        } else {
#pragma unroll
          for (int i = 0; i < LE; i++) xg[i] = a.x[ci[i]];
        }
        run_sums(vv, xg, fb, p);
      }
      // Piece 0 is the run before the first inner start bit; the sum the
      // lane leaves open is its 8th slot's (a lane holding fewer than 8
      // elements ends the unit and closes its last segment).
      const uint32_t inner = fb & ~1u;
      const int f1 = inner ? __builtin_ctz(inner) : nv;
      const int has_close = nv > 0 && (inner != 0 || tclosed);
      const double run = nv > 0 && !tclosed ? p[LE - 1] : 0.0;
      const int cnt = __builtin_popcount(fb);
      // segmented inclusive scan c(t) = run(t) + (has_close(t) ? 0 : c(t-1)),
      // and an inclusive scan of the start-bit counts, across the wave
      double v = run;
      int sg = has_close, cn = cnt;
      scan_step<0x111, 0xf, 0xf>(v, sg, cn);  // row_shr:1
      scan_step<0x112, 0xf, 0xf>(v, sg, cn);  // row_shr:2
      scan_step<0x114, 0xf, 0xe>(v, sg, cn);  // row_shr:4
      scan_step<0x118, 0xf, 0xc>(v, sg, cn);  // row_shr:8
      scan_step<0x142, 0xa, 0xf>(v, sg, cn);  // row_bcast:15 into rows 1, 3
      scan_step<0x143, 0xc, 0xf>(v, sg, cn);  // row_bcast:31 into rows 2, 3
      if (lane == kWave - 1) {
        s_wv[par][wid] = v;
        s_ws[par][wid] = sg;
        s_wc[par][wid] = cn;
      }
      __syncthreads();
      double prev = carry, tot = carry;
      int cb = base, totc = base;
#pragma unroll
      for (int w = 0; w < NW; w++) {
        const double wv = s_wv[par][w];
        const int ws = s_ws[par][w], wc = s_wc[par][w];
        if (w < wid) {
          prev = wv + (ws ? 0.0 : prev);
          cb += wc;
        }
        tot = wv + (ws ? 0.0 : tot);
        totc += wc;
      }
      const double c_incl = v + (sg ? 0.0 : prev);
      double carry_in = __shfl_up(c_incl, 1, kWave);
      if (lane == 0) carry_in = prev;
      const int excl = cb + cn - cnt;  // output of the first segment begun by this thread
      // The sums closed in this chunk have outputs in [base - 1, totc): they
      // go to LDS and leave in one coalesced pass, unless the chunk closes
      // more than NOUT of them (then each is stored where it closes).
      const int kb = base - 1;
      const bool via_lds = totc - kb <= NOUT;  // uniform
      auto put = [&](int k, double v) {
        if (via_lds) {
          s_out[k - kb] = v;
          s_ok[k - kb] = 1;
        } else {
          seg_emit(a, part, k, v);
        }
      };
      // Every element that ends a segment emits its running sum (plus the
      // carry from earlier threads for piece 0) to the segment's output.
      // ends: bit i set when element i is the last of its segment.
      const uint32_t ends = nv > 0 ? ((fb >> 1) | ((uint32_t)tclosed << (nv - 1))) & ((1u << nv) - 1) : 0u;
      int k = excl - 1;  // the output of the segment holding element i
#pragma unroll
      for (int i = 0; i < LE; i++) {
        k += (fb >> i) & 1;
        if ((ends >> i) & 1) put(k, i < f1 ? p[i] + carry_in : p[i]);
      }
      if (via_lds) {
        // s_out is written only after the next chunk's barrier, which every
        // thread reaches after this pass: one buffer suffices
        __syncthreads();
        for (int t = tid; t < totc - kb; t += kSegBlock) {
          if (s_ok[t]) {
            seg_emit(a, part, kb + t, s_out[t]);
            s_ok[t] = 0;
          }
        }
      }
      carry = tot;
      base = totc;
    }
  }
}

// Kernel 4 with column bands: wave w owns rows 64w .. 64w+63.  Row r's
// slice in band b sits at base[w * 32 + b] plus the number of the wave's
// earlier rows with a slice in band b (a ballot and mbcnt), so for each band
// the wave reads one contiguous run of partials.  Every band's load is
// issued before the first is added (a lane without a slice in a band reads
// the slot its next row's slice would take, within the same lines, and
// drops it).  A row adds its slices in band order; long rows (mask
// bit 31) are left to the long-row fold, rows without nonzeros get beta * y.
// The waves are persistent (a grid-stride walk over the 64-row groups):
// one short wave per group is bound by the rate at which waves launch.
constexpr int kMaxBandsDev = 31;  // bands per plan at most (bit 31 of a row's mask marks a long row)

__global__ __launch_bounds__(kBlock) void k_band_reduce(const uint32_t *__restrict__ mask,
                                                        const int32_t *__restrict__ base,
                                                        int nb, const double *__restrict__ partial, int last,
                                                        int m, double alpha, double beta, double *__restrict__ y) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x % kWave;
  const int nw = (m + kWave - 1) / kWave;
  const int step = __builtin_amdgcn_readfirstlane((int)(gridDim.x * (blockDim.x / kWave)));
  for (int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave)); w < nw;
       w += step) {
    const int64_t r = (int64_t)w * kWave + lane;
    const uint32_t mk = r < m ? __builtin_nontemporal_load(&mask[r]) : 1u << 31;
    // the wave's 32 slot bases in one go (128 bytes, 16-byte aligned)
    int32_t bw[kMaxBandsDev + 1];
    {
      const vi4 *p4 = reinterpret_cast<const vi4 *>(base + (int64_t)w * (kMaxBandsDev + 1));
#pragma unroll
      for (int q = 0; q < (kMaxBandsDev + 1) / 4; q++) {
        const vi4 t = p4[q];
        bw[4 * q] = t.x;
        bw[4 * q + 1] = t.y;
        bw[4 * q + 2] = t.z;
        bw[4 * q + 3] = t.w;
      }
    }
    double v[kMaxBandsDev];
#pragma unroll
    for (int b = 0; b < kMaxBandsDev; b++) {
      const uint32_t bit = (mk >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      const int below =
          (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0));
      const int at = bw[b] + below;
      v[b] = b < nb ? __builtin_nontemporal_load(&partial[at < last ? at : last]) : 0.0;
    }
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < kMaxBandsDev; b++)
      if ((mk >> b) & 1u) s += v[b];
    if (r < m && !(mk >> 31))
      y[r] = mk == 0 ? (beta == 0.0 ? 0.0 : beta * y[r]) : axpby(alpha, s, beta, beta == 0.0 ? 0.0 : y[r]);
  }
}

hipError_t launch_band_reduce(const uint32_t *mask, const int32_t *base, int nb, const double *partial,
                              int64_t nslots, int m, double alpha, double beta, double *y, hipStream_t st) {
  if (m <= 0 || nb <= 0 || nslots <= 0) return hipSuccess;
  if (nb > kMaxBandsDev) return hipErrorInvalidValue;
  const int64_t waves = ((int64_t)m + kWave - 1) / kWave;
  const int64_t blocks = std::min<int64_t>((waves * kWave + kBlock - 1) / kBlock, 2048);
  hipLaunchKernelGGL(k_band_reduce, dim3((unsigned)blocks), dim3(kBlock), 0, st, mask, base, nb, partial,
                     (int)(nslots - 1), m, alpha, beta, y);
  return hipGetLastError();
}

template <typename K>
static int seg_cap(K kern) {
  int nb = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, kSegBlock, 0) != hipSuccess || nb < 1) nb = 1;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  const int cap = nb * cus / kXcd * kXcd;
  return cap < kXcd ? kXcd : cap;
}

template <typename V, typename X, bool STAGE, int NOUT>
static hipError_t seg_launch(const SegArgs<V, X> &a, hipStream_t st) {
  if (a.nunits <= 0) return hipSuccess;
  static const int cap = seg_cap(k_seg<V, X, STAGE, NOUT>);
  int g = direct_grid(a.nunits);
  if (g > cap) g = cap;
  hipLaunchKernelGGL((k_seg<V, X, STAGE, NOUT>), dim3(g), dim3(kSegBlock), 0, st, a);
  return hipGetLastError();
}

// `wide_out`: units of one- or two-element segments (column-band slices)
// close up to a whole chunk of sums, and gathering all of them in LDS
// (kSegOutWide) beats storing each where it closes; only kernels without
// the x window have the LDS for it.
template <typename V>
static hipError_t seg_launch_any(const SegArgs<V, double> &a, bool stage, bool wide_out, hipStream_t st) {
  if (stage) return seg_launch<V, double, true, kSegOut>(a, st);
  return wide_out ? seg_launch<V, double, false, kSegOutWide>(a, st) : seg_launch<V, double, false, kSegOut>(a, st);
}

hipError_t launch_seg_f64(const SegArgs<double, double> &a, bool stage, bool wide_out, hipStream_t st) {
  return seg_launch_any(a, stage, wide_out, st);
}
hipError_t launch_seg_f32(const SegArgs<float, double> &a, bool stage, bool wide_out, hipStream_t st) {
  return seg_launch_any(a, stage, wide_out, st);
}

}  // namespace spx
