// kernels.hip -- hand-written gfx950 (CDNA4, wave64) kernels for
// y = alpha*A*x + beta*y on a CSR matrix, plus the device-side partitioner
// and the CSR5-style tile conversion.
//
// Work is cut into partitions (nnz-balanced or row blocks) and each partition
// into workgroup "units".  A row whose nonzeros fall in more than one unit is
// never written by a unit: every unit writes its piece of such a row into
// one of its two carry slots, tagged as the row's first piece, its final
// piece or a piece between, and k_slot_fix sums each row's pieces in slot
// order (nonzero order) and applies alpha and the beta term once.  Rows a
// unit holds whole are written directly.  Summation order is fixed, so
// results are bitwise reproducible; no atomics touch y.
//
// beta == 0 never reads y (BLAS convention), so y may hold NaN on input.
#include "seg_device.h"

namespace spx {

// ---- helpers -----------------------------------------------------------------

// upper_bound(rowptr[lo..hi), k) - 1 over absolute row indices.
__device__ inline int row_search(const int32_t *rowptr, int lo, int hi, int64_t k) {
  while (lo < hi) {
    int mid = lo + ((hi - lo) >> 1);
    if ((int64_t)rowptr[mid] <= k)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - 1;
}

__device__ inline double y_update(double alpha, double beta, double sum, const double *y, int r) {
  return beta == 0.0 ? alpha * sum : alpha * sum + beta * y[r];
}

// A final piece sums its row: back over the row's middle pieces (and unused
// slots) to its first piece, then forward in slot order.
__device__ inline void slot_fix(const int32_t *__restrict__ crow, const double *__restrict__ cval, int64_t k,
                                double alpha, double beta, double *__restrict__ y) {
  const int32_t c = crow[k];
  if (c == -1 || !((uint32_t)c & kSlotFinal)) return;
  const int32_t r = c & kSlotRow;
  int64_t j = k - 1;
  while (j >= 0 && crow[j] != (r | kSlotFirst)) j--;
  double s = 0.0;
  for (int64_t i = j < 0 ? 0 : j; i < k; i++)
    if (crow[i] != -1) s += cval[i];
  s += cval[k];
  y[r] = y_update(alpha, beta, s, y, r);
}

// Index of the last partition whose `off` member is <= g (parts sorted by off).
template <typename F>
__device__ inline int find_part(const Part *parts, int nparts, int64_t g, F off) {
  int lo = 0, hi = nparts;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (off(parts[mid]) <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Row pointer of partition q seen as its own CSR matrix (the local row
// pointer of SB/src/dspmv_mgpu_v1.cu:125-133, with the split rows clamped to
// the partition): lrp(j) = clamp(rowptr[row_lo + j] - s, 0, nnz_q).
__device__ inline int64_t lrp(const int32_t *rowptr, const Part &q, int64_t j) {
  int64_t v = (int64_t)rowptr[q.row_lo + j] - q.s;
  int64_t nz = q.e - q.s + 1;
  return v < 0 ? 0 : (v > nz ? nz : v);
}

// upper_bound over lrp(j), j in [lo, hi), minus 1.
__device__ inline int64_t local_row_search(const int32_t *rowptr, const Part &q, int64_t lo,
                                           int64_t hi, int64_t k) {
  while (lo < hi) {
    int64_t mid = lo + ((hi - lo) >> 1);
    if (lrp(rowptr, q, mid) <= k)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - 1;
}

__device__ inline void put_carry(int32_t *crow, double *cval, int64_t slot, int row, double v) {
  crow[slot] = row;
  cval[slot] = v;
}

// ---- partitioner ---------------------------------------------------------------

__global__ void k_partition(const int32_t *__restrict__ rowptr, int m, int64_t nnz, int P, int kind,
                            Part *__restrict__ parts) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  Part q = {};
  if (kind == PART_ROWS) {
    // SB/src/dspmv_mgpu_baseline.cu:64-65
    int r0 = (int)((int64_t)p * m / P), r1 = (int)((int64_t)(p + 1) * m / P) - 1;
    q.r0 = r0;
    q.r1 = r1;
    q.s = rowptr[r0];
    q.e = (int64_t)rowptr[r1 + 1] - 1;
    q.row_lo = r0;
    q.row_hi = r1 + 1;
  } else {
    int64_t s, e, sn;
    if (kind == PART_NNZ) {
      // SB/src/dspmv_mgpu_v1.cu:62-69: floor of a double quotient
      s = (int64_t)floor((double)((int64_t)p * nnz) / P);
      e = (int64_t)floor((double)((int64_t)(p + 1) * nnz) / P) - 1;
    } else {
      // SB/src/dspmv_mgpu_v2.cu:229-236: integer quotient
      s = (int64_t)p * nnz / P;
      e = (int64_t)(p + 1) * nnz / P - 1;
    }
    sn = e + 1;
    q.s = s;
    q.e = e;
    q.r0 = row_search(rowptr, 0, m + 1, s);
    q.r1 = row_search(rowptr, 0, m + 1, e);
    q.sflag = s > rowptr[q.r0];
    q.eflag = e < (int64_t)rowptr[q.r1 + 1] - 1;
    q.row_lo = p == 0 ? 0 : q.r0;
    if (p == P - 1)
      q.row_hi = m;
    else
      q.row_hi = q.eflag ? q.r1 + 1 : row_search(rowptr, 0, m + 1, sn);
  }
  parts[p] = q;
}

__global__ void k_merge_coords(const int32_t *__restrict__ rowptr, const Part *__restrict__ parts,
                               int nparts, int64_t total, int tw, Coord *__restrict__ coords) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int pi = find_part(parts, nparts, g, [](const Part &q) { return q.coord_off; });
  const Part q = parts[pi];
  int64_t u = g - q.coord_off;
  int64_t nrows = (int64_t)q.row_hi - q.eflag - q.row_lo;  // row-end items
  int64_t nnzs = q.e + 1 - q.s;
  int64_t L = nrows + nnzs;
  int64_t d = u * tw;
  if (d > L) d = L;
  // merge-path search: a row end i precedes nonzero k when rowptr[i+1] <= k
  int64_t lo = d - nnzs > 0 ? d - nnzs : 0, hi = d < nrows ? d : nrows;
  while (lo < hi) {
    int64_t piv = (lo + hi) >> 1;
    if ((int64_t)rowptr[q.row_lo + piv + 1] <= q.s + (d - piv - 1))
      lo = piv + 1;
    else
      hi = piv;
  }
  coords[g].row = (int32_t)(q.row_lo + lo);
  coords[g].nz = (int32_t)(q.s + d - lo);
}

// ---- CSR5-style tile conversion (64 lanes x sigma) -------------------------------
// Each partition is converted as its own CSR matrix (local rows
// [0, row_hi - row_lo), local nonzeros [0, e - s]).  The layout follows
// SB/include/detail/cuda/format_cuda.h: tile pointer = upper_bound - 1 with
// bit 31 flagging tiles that cover an empty row (21-95), one row-start bit per
// element (140-157), and the per-tile transpose to lane-fastest order
// (525-585), with omega = 64 lanes.

__global__ void k_tile_ptr(const int32_t *__restrict__ rowptr, const Part *__restrict__ parts,
                           int nparts, int64_t total, int sigma, uint32_t *__restrict__ tp) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int pi = find_part(parts, nparts, g, [](const Part &q) { return q.tp_off; });
  const Part q = parts[pi];
  int64_t t = g - q.tp_off;
  int64_t nz = q.e - q.s + 1;
  int64_t pos = t * kOmega * sigma;
  if (pos > nz) pos = nz;
  int64_t nrows = q.row_hi - q.row_lo;
  tp[g] = (uint32_t)local_row_search(rowptr, q, 0, nrows + 1, pos);
}

__global__ void k_tile_dirty(const int32_t *__restrict__ rowptr, const Part *__restrict__ parts,
                             int nparts, int64_t total_tiles, uint32_t *__restrict__ tp) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total_tiles) return;
  int pi = find_part(parts, nparts, g, [](const Part &q) { return q.tile_off; });
  const Part q = parts[pi];
  int64_t t = g - q.tile_off;
  uint32_t start = tp[q.tp_off + t] & 0x7FFFFFFFu;
  uint32_t stop = tp[q.tp_off + t + 1] & 0x7FFFFFFFu;
  if (start == stop) return;
  for (uint32_t r = start; r < stop; r++) {
    if (lrp(rowptr, q, r) == lrp(rowptr, q, r + 1)) {
      tp[q.tp_off + t] = start | 0x80000000u;
      return;
    }
  }
}

__global__ void k_tile_flags(const int32_t *__restrict__ rowptr, const Part *__restrict__ parts,
                             int nparts, int64_t total_words, int sigma,
                             uint32_t *__restrict__ flags) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total_words) return;
  int64_t gt = g / kOmega;
  int lane = (int)(g % kOmega);
  int pi = find_part(parts, nparts, gt, [](const Part &q) { return q.tile_off; });
  const Part q = parts[pi];
  int64_t t = gt - q.tile_off;
  int64_t nz = q.e - q.s + 1;
  int64_t nrows = q.row_hi - q.row_lo;
  int64_t b = t * kOmega * sigma + (int64_t)lane * sigma;
  // first local row whose start is >= b
  int64_t lo = 0, hi = nrows;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (lrp(rowptr, q, mid) < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  uint32_t w = 0;
  for (int64_t r = lo; r < nrows; r++) {
    int64_t st = lrp(rowptr, q, r);
    if (st >= b + sigma || st >= nz) break;
    if (st < lrp(rowptr, q, r + 1)) w |= 1u << (uint32_t)(st - b);
  }
  flags[g] = w;
}

template <typename VO>
__global__ void k_tile_transpose(const Part *__restrict__ parts, int nparts, int64_t total, int sigma,
                                 const int32_t *__restrict__ col, const double *__restrict__ val,
                                 int32_t *__restrict__ colt, VO *__restrict__ valt) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int64_t w = (int64_t)kOmega * sigma;
  int64_t gt = g / w;
  int64_t within = g - gt * w;
  int k = (int)(within / kOmega), lane = (int)(within % kOmega);
  int pi = find_part(parts, nparts, gt, [](const Part &q) { return q.tile_off; });
  const Part q = parts[pi];
  int64_t t = gt - q.tile_off;
  int64_t pos = t * w + (int64_t)lane * sigma + k;
  int64_t nz = q.e - q.s + 1;
  if (pos < nz) {
    colt[g] = col[q.s + pos];
    valt[g] = (VO)val[q.s + pos];
  } else {
    colt[g] = 0;
    valt[g] = (VO)0;
  }
}

// ---- small kernels -----------------------------------------------------------------

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_to_f32(const double *__restrict__ v, float *__restrict__ o, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = (float)v[i];
}

__global__ void k_empty_rows(const int32_t *__restrict__ rows, int64_t n, double beta,
                             double *__restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    int r = rows[i];
    y[r] = beta == 0.0 ? 0.0 : beta * y[r];
  }
}

__global__ void k_find_empty(const int32_t *__restrict__ rowptr, int m, int32_t *__restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) flag[i] = rowptr[i] == rowptr[i + 1];
}

__global__ void k_reset(int *q, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) q[i] = 0;
}

// ---- long rows: the fold of the per-column-block partial sums (long_reduce, seg_device.h)

__global__ __launch_bounds__(kBlock) void k_long_reduce(const double *__restrict__ partial, int nlong,
                                                        const int32_t *__restrict__ lsp,
                                                        const int32_t *__restrict__ lseg,
                                                        const int32_t *__restrict__ rows, double alpha,
                                                        double beta, int assign, double *__restrict__ y) {
  if ((int)blockIdx.x < nlong) long_reduce(partial, lsp, lseg, rows, blockIdx.x, alpha, beta, assign, y);
}

// The end of an execute of kernels 1-3 in one launch: workgroups below nlong
// fold a long row each, the rest run the slot fix-up (disjoint rows).
__global__ __launch_bounds__(kBlock) void k_finish(const double *__restrict__ partial, int nlong,
                                                   const int32_t *__restrict__ lsp, const int32_t *__restrict__ lseg,
                                                   const int32_t *__restrict__ rows, int assign,
                                                   const int32_t *__restrict__ crow, const double *__restrict__ cval,
                                                   int64_t nslots, double alpha, double beta,
                                                   double *__restrict__ y) {
  if ((int)blockIdx.x < nlong) {
    long_reduce(partial, lsp, lseg, rows, blockIdx.x, alpha, beta, assign, y);
  } else {
    const int64_t k = (int64_t)(blockIdx.x - nlong) * kBlock + threadIdx.x;
    if (k < nslots) slot_fix(crow, cval, k, alpha, beta, y);
  }
}

// ---- kernel 1: CSR-vector, L lanes per row --------------------------------------
// Bell & Garland's CSR-vector for wave64: a group of L lanes sums a row.  A
// unit is up to kVecUnitRows consecutive rows of one partition holding at
// most kBlock*kVecItems nonzeros (the plan cuts units by nonzeros, so units
// fill about the same loads whatever the row lengths; a small matrix gets
// smaller units, enough to spread over the CUs).  The unit's nonzeros are
// loaded with coalesced loads that need no row bounds, so they are requested
// at once with the x window and the unit's row pointer, both copied to LDS
// by the load units; the products go to LDS and each row's L lanes add
// theirs from there, G = kBlock/L rows at a time.  A row of more than
// kBlock*kVecItems nonzeros is a unit alone: the whole workgroup sums it
// with coalesced loads.  The group's sum is reduced with DPP lane
// permutations.  The partition's split rows go to its carry slots.

// One DPP permutation of a double (all lanes have a source for these controls).
template <int CTRL>
__device__ inline double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Sum over each aligned group of L lanes; every lane of the group gets it.
template <int L>
__device__ inline double group_sum(double s) {
  if (L >= 2) s += dpp_d<0xb1>(s);   // quad_perm [1,0,3,2]
  if (L >= 4) s += dpp_d<0x4e>(s);   // quad_perm [2,3,0,1]
  if (L >= 8) s += dpp_d<0x141>(s);  // row_half_mirror: the other quad of 8
  if (L >= 16) s += dpp_d<0x140>(s); // row_mirror: the other 8 of 16
  if (L >= 32) s += __shfl_xor(s, 16, kWave);
  if (L >= 64) s += __shfl_xor(s, 32, kWave);
  return s;
}

// Kernel 1's result for local row lr of unit gu: y, or, for a row the
// partition's bounds cut (begins: the row starts inside the partition, ends:
// it ends inside it), the unit's carry slot 2gu (first row) or 2gu + 1.
template <typename V>
__device__ inline void vec_put(const KArgs<V> &a, int64_t gu, int r_lo, int lr, bool begins, bool ends, double t) {
  const int r = r_lo + lr;
  if (begins && ends) {
    a.y[r] = y_update(a.alpha, a.beta, t, a.y, r);
  } else {
    const int64_t slot = 2 * gu + (lr == 0 ? 0 : 1);
    a.crow[slot] = (int32_t)((uint32_t)r | (begins ? (uint32_t)kSlotFirst : 0u) | (ends ? kSlotFinal : 0u));
    a.cval[slot] = t;
  }
}

template <typename V, int L>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_vector(KArgs<V> a) {
  constexpr int G = kBlock / L;
  constexpr int PCAP = kBlock * kVecItems;
  constexpr int XS = (kSegStage + 2) > PCAP ? (kSegStage + 2) : PCAP;
  // the x window, then (after every gather) the unit's products
  __shared__ __attribute__((aligned(16))) double s_x[XS];
  __shared__ int32_t s_rp[kVecUnitRows + 1];  // the unit's row pointer
  __shared__ double s_red[kBlock / kWave];
  __shared__ int s_slot;
  const int tid = threadIdx.x, g = tid / L, l = tid % L, lane = tid % kWave, wid = tid / kWave;
  // rows differ in length, so without the queue the XCDs take unit ranges
  // of about equal nonzeros (vxr) rather than equal unit counts
  int64_t iter = 0;
  auto next = [&]() {
    return a.vxr && !a.queue ? next_unit_ranged(a.vxr, iter++, a.unit_count)
                             : next_unit(a.queue, iter++, a.unit_count, &s_slot);
  };
  int64_t u = next();
  int dv = desc_prefetch(a.vunits + a.unit_begin, u, a.unit_count, lane);
  while (u < a.unit_count) {
    const int dcur = dv;
    const int64_t gu = a.unit_begin + u;
    u = next();
    dv = desc_prefetch(a.vunits + a.unit_begin, u, a.unit_count, lane);
    const int r_lo = desc_field(dcur, 0), r_hi = desc_field(dcur, 1), s0 = desc_field(dcur, 2),
              e1 = desc_field(dcur, 3), cmin = desc_field(dcur, 4), cwin = desc_field(dcur, 5),
              z0 = desc_field(dcur, 6), z1 = desc_field(dcur, 7);
    const int nr = r_hi - r_lo, nzu = z1 - z0;
    __syncthreads();  // the previous unit is done with s_x, s_rp and s_red
    if (nzu <= PCAP) {
      const bool stg = cwin > 0;
      int lead = 0;
      if (stg) lead = stage_window<kBlock>(a.x, cmin, cwin, s_x, tid);
      // the unit's row pointer rowptr[r_lo .. r_hi] into s_rp
      for (int q0 = wid * kWave; q0 <= nr; q0 += kBlock)
        if (q0 + lane <= nr)
          __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(a.rowptr + r_lo + q0 + lane),
                                           (__attribute__((address_space(3))) void *)(s_rp + q0), 4, 0, 0);
      int ci[kVecItems];
      V vv[kVecItems];
      __builtin_amdgcn_s_setprio(2);
#pragma unroll
      for (int i = 0; i < kVecItems; i++) {
        const int k = tid + i * kBlock;
        ci[i] = k < nzu ? __builtin_nontemporal_load(a.col + z0 + k) : 0;
        vv[i] = k < nzu ? __builtin_nontemporal_load(a.val + z0 + k) : (V)0;
      }
      __builtin_amdgcn_s_setprio(0);
      stage_wait();
      __syncthreads();  // the window and the row pointer are in LDS
      double pv[kVecItems];
      if (stg) {
#pragma unroll
        for (int i = 0; i < kVecItems; i++)
          pv[i] = prod(vv[i], s_x[lead + (tid + i * kBlock < nzu ? ci[i] - cmin : 0)]);
        __syncthreads();  // every gather is done: the products overwrite the window
      } else {
#pragma unroll
        for (int i = 0; i < kVecItems; i++) pv[i] = prod(vv[i], a.x[ci[i]]);
      }
#pragma unroll
      for (int i = 0; i < kVecItems; i++) {
        const int k = tid + i * kBlock;
        if (k < nzu) s_x[k] = pv[i];
      }
      __syncthreads();  // the unit's products are in LDS
      for (int k0 = 0; k0 < nr; k0 += G) {
        const int lr = k0 + g;
        double t = 0.0;
        int rb = 0, re = 0;
        if (lr < nr) {
          rb = s_rp[lr];
          re = s_rp[lr + 1];
          const int b = rb > s0 ? rb : s0, e = re < e1 ? re : e1;
          for (int j = b - z0 + l; j < e - z0; j += L) t += s_x[j];
        }
        t = group_sum<L>(t);
        if (l == 0 && lr < nr) vec_put(a, gu, r_lo, lr, rb >= s0, re <= e1, t);
      }
    } else {
      // one row longer than the unit's load capacity: every thread takes
      // every kBlock-th nonzero, gathering from x
      const int rb = a.rowptr[r_lo], re = a.rowptr[r_lo + 1];
      const int b = rb > s0 ? rb : s0, e = re < e1 ? re : e1;
      double t = 0.0;
      int k = b + tid;
      for (; k + 3 * kBlock < e; k += 4 * kBlock) {
        int c[4];
        V v[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          c[i] = __builtin_nontemporal_load(a.col + k + i * kBlock);
          v[i] = __builtin_nontemporal_load(a.val + k + i * kBlock);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) t += prod(v[i], a.x[c[i]]);
      }
      for (; k < e; k += kBlock) t += prod(a.val[k], a.x[a.col[k]]);
      t = group_sum<kWave>(t);
      if (lane == 0) s_red[wid] = t;
      __syncthreads();
      if (tid == 0) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; w++) s += s_red[w];
        vec_put(a, gu, r_lo, 0, rb >= s0, re <= e1, s);
      }
    }
  }
}

// ---- kernel 2: merge path ------------------------------------------------------
// Merrill & Garland (SC'16): the merge of row ends and nonzero indices is cut
// into equal pieces of kBlock*IPT items, one per unit (MUnit: the merge-path
// coordinates at both ends, found at plan time), so every unit does the same
// work however the rows are shaped.  Inside a unit the row ends are staged in
// LDS and turned into start bits (a row end inside the unit starts the next
// nonempty row's run of nonzeros).  The products val*x[col] are formed with
// coalesced loads (x gathered from the LDS window when the unit's columns fit
// one) and transposed through LDS so that each thread owns IPT consecutive
// nonzeros; running sums restart at the start bits, a block-wide segmented
// scan (DPP steps across each wave, wave totals through LDS) joins the runs
// that cross threads, and the scanned values go back to LDS, where each row
// reads its sum at its last nonzero and leaves in one coalesced pass.  A row
// cut by a unit boundary leaves pieces in the units' carry slots (2u: the
// piece continuing a row begun earlier, 2u+1: the piece open at the unit's
// end), tagged with kSlotFirst / kSlotFinal, and k_finish adds them.  Every
// slot is rewritten each time, so no slot state carries over between calls.

template <int CTRL, int ROWS, int BANKS>
__device__ inline void scan_step_vs(double &v, int &sg) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, BANKS, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, BANKS, false);
  const int su = __builtin_amdgcn_update_dpp(0, sg, CTRL, ROWS, BANKS, false);
  if (!sg) {
    v += __hiloint2double(hi, lo);
    sg |= su;
  }
}

template <typename V, int IPT>
__global__ __launch_bounds__(kBlock) void k_merge(KArgs<V> a) {
  static_assert(IPT >= 1 && IPT <= 32, "at most 32 nonzeros per thread");
  constexpr int TW = kBlock * IPT;
  constexpr int NW = kBlock / kWave;
  // the x window, then (after every gather) the unit's products, then their
  // scan; product k sits at k + k/8 (one pad per 8), so a thread's IPT
  // consecutive products and its neighbours' fall in different LDS banks
  constexpr int XS = (kSegStage + 2) > TW + TW / 8 ? (kSegStage + 2) : TW + TW / 8;
  __shared__ __attribute__((aligned(16))) double s_x[XS];
  __shared__ int32_t s_re[TW + 1];          // rowptr[r0 + k], k = 0..rows
  __shared__ uint32_t s_fl[TW / 32 + 2];    // start bits, one per nonzero of the unit
  __shared__ double s_wv[NW];
  __shared__ int s_ws[NW];
  __shared__ int s_slot;
  double *const s_p = s_x;
  const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
  // the start bits are cleared here and again after each unit has read them
  for (int w = tid; w < TW / 32 + 2; w += kBlock) s_fl[w] = 0u;
  __syncthreads();
  int64_t iter = 0;
  int64_t u = next_unit(a.queue, iter++, a.unit_count, &s_slot);
  int dv = desc_prefetch(a.munits + a.unit_begin, u, a.unit_count, lane);
  while (u < a.unit_count) {
    const int dcur = dv;
    const int64_t gu = a.unit_begin + u;
    u = next_unit(a.queue, iter++, a.unit_count, &s_slot);
    dv = desc_prefetch(a.munits + a.unit_begin, u, a.unit_count, lane);
    const int i0 = desc_field(dcur, 0), j0 = desc_field(dcur, 1), r1 = desc_field(dcur, 2), j1 = desc_field(dcur, 3);
    const int cmin = desc_field(dcur, 4), cwin = desc_field(dcur, 5);
    const int nrows = r1 - i0, nnzs = j1 - j0;
    const bool stg = cwin > 0;
    int lead = 0;
    // the x window, the unit's nonzeros and its row ends are all requested
    // before the first wait
    if (stg) lead = stage_window<kBlock>(a.x, cmin, cwin, s_x, tid);
    int ci[IPT];
    V vv[IPT];
    __builtin_amdgcn_s_setprio(2);
#pragma unroll
    for (int r = 0; r < IPT; r++) {
      const int k = tid + r * kBlock;
      ci[r] = k < nnzs ? __builtin_nontemporal_load(a.col + j0 + k) : 0;
      vv[r] = k < nnzs ? __builtin_nontemporal_load(a.val + j0 + k) : (V)0;
    }
    __builtin_amdgcn_s_setprio(0);
    // a row end at unit position p < nnzs: nonzero p starts a row
    for (int k = tid; k <= nrows; k += kBlock) {
      const int re = a.rowptr[i0 + k];
      s_re[k] = re;
      const int p = re - j0;
      if (k > 0 && p >= 0 && p < nnzs) atomicOr(&s_fl[p >> 5], 1u << (p & 31));
    }
    double pv[IPT];
    if (stg) {
      stage_wait();
      __syncthreads();  // the window is in LDS
#pragma unroll
      for (int r = 0; r < IPT; r++) pv[r] = prod(vv[r], s_x[lead + (tid + r * kBlock < nnzs ? ci[r] - cmin : 0)]);
      __syncthreads();  // every gather is done: the products overwrite the window
    } else {
#pragma unroll
      for (int r = 0; r < IPT; r++) pv[r] = prod(vv[r], a.x[ci[r]]);
    }
#pragma unroll
    for (int r = 0; r < IPT; r++) {
      const int k = tid + r * kBlock;
      if (k < nnzs) s_p[k + (k >> 3)] = pv[r];
    }
    __syncthreads();  // products and start bits are in LDS

    // this thread's nonzeros [q0, q0 + IPT) of the unit, and their start bits
    const int q0 = tid * IPT;
    double p[IPT];
#pragma unroll
    for (int k = 0; k < IPT; k++) p[k] = q0 + k < nnzs ? s_p[q0 + k + ((q0 + k) >> 3)] : 0.0;
    const uint64_t w2 = (uint64_t)s_fl[q0 >> 5] | ((uint64_t)s_fl[(q0 >> 5) + 1] << 32);
    const uint32_t fb = (uint32_t)(w2 >> (q0 & 31)) & (IPT == 32 ? 0xFFFFFFFFu : ((1u << IPT) - 1u));
    // running sums that restart at every start bit
#pragma unroll
    for (int k = 1; k < IPT; k++)
      if (!((fb >> k) & 1u)) p[k] = p[k - 1] + p[k];

    // segmented inclusive scan c(t) = run(t) + (t holds a start bit ? 0 : c(t-1))
    double v = p[IPT - 1];
    int sg = fb != 0u;
    scan_step_vs<0x111, 0xf, 0xf>(v, sg);  // row_shr:1
    scan_step_vs<0x112, 0xf, 0xf>(v, sg);  // row_shr:2
    scan_step_vs<0x114, 0xf, 0xe>(v, sg);  // row_shr:4
    scan_step_vs<0x118, 0xf, 0xc>(v, sg);  // row_shr:8
    scan_step_vs<0x142, 0xa, 0xf>(v, sg);  // row_bcast:15 into rows 1, 3
    scan_step_vs<0x143, 0xc, 0xf>(v, sg);  // row_bcast:31 into rows 2, 3
    if (lane == kWave - 1) {
      s_wv[wid] = v;
      s_ws[wid] = sg;
    }
    __syncthreads();  // also: every thread has read its start bits
    for (int w = tid; w < TW / 32 + 2; w += kBlock) s_fl[w] = 0u;
    double prev = 0.0;
#pragma unroll
    for (int w = 0; w < NW; w++)
      if (w < wid) prev = s_wv[w] + (s_ws[w] ? 0.0 : prev);
    const double c_incl = v + (sg ? 0.0 : prev);
    double carry_in = __shfl_up(c_incl, 1, kWave);
    if (lane == 0) carry_in = prev;
    // the run open at the thread's start continues the previous thread's
    const int first = fb ? __builtin_ctz(fb) : IPT;
#pragma unroll
    for (int k = 0; k < IPT; k++) {
      if (k < first) p[k] += carry_in;
      if (q0 + k < nnzs) s_p[q0 + k + ((q0 + k) >> 3)] = p[k];
    }
    __syncthreads();  // the scanned values are in LDS

    // row k's sum is the scanned value at its last nonzero in the unit
    const int rb = s_re[0];
    const bool head_split = nrows > 0 && rb < j0;  // row r0 began in an earlier unit and ends here
    for (int k = tid; k < nrows; k += kBlock) {
      const int b0 = s_re[k] - j0, e = s_re[k + 1] - j0;
      const double sv = e > (b0 > 0 ? b0 : 0) ? s_p[(e - 1) + ((e - 1) >> 3)] : 0.0;
      if (k == 0 && head_split) {
        a.crow[2 * gu] = (int32_t)((uint32_t)i0 | kSlotFinal);
        a.cval[2 * gu] = sv;
      } else {
        a.y[i0 + k] = y_update(a.alpha, a.beta, sv, a.y, i0 + k);
      }
    }
    if (tid == kBlock - 1) {
      // nonzeros after the unit's last row end belong to row r1, still open
      int32_t hc = -1, tc = -1;
      double hv = 0.0, tv = 0.0;
      const double last = nnzs > 0 ? s_p[(nnzs - 1) + ((nnzs - 1) >> 3)] : 0.0;
      if (nrows == 0) {
        if (nnzs > 0) {
          if (rb >= j0) {
            tc = i0 | kSlotFirst;  // row r0 = r1 begins in this unit
            tv = last;
          } else {
            hc = i0;  // a middle piece of a row spanning this unit
            hv = last;
          }
        }
      } else if (j1 > s_re[nrows]) {
        tc = r1 | kSlotFirst;
        tv = last;
      }
      if (!head_split) {
        a.crow[2 * gu] = hc;
        a.cval[2 * gu] = hv;
      }
      a.crow[2 * gu + 1] = tc;
      a.cval[2 * gu + 1] = tv;
    }
    __syncthreads();  // LDS is rewritten by the next unit
  }
}

__global__ void k_slot_fix(const int32_t *__restrict__ crow, const double *__restrict__ cval, int64_t nslots,
                           double alpha, double beta, double *__restrict__ y) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nslots) slot_fix(crow, cval, k, alpha, beta, y);
}

__global__ void k_merge_units(const Coord *__restrict__ coords, int64_t units, MUnit *__restrict__ out) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= units) return;
  const Coord c0 = coords[u], c1 = coords[u + 1];
  out[u] = MUnit{c0.row, c0.nz, c1.row, c1.nz, c0.nz, c1.nz, 0, 0};  // window fields: the nonzero range
}

// ---- kernel 3: CSR5-style tiles -------------------------------------------------
// Liu & Vinter's CSR5 (ICS'15) recast for wave64.  Tiles of 64 lanes x sigma
// nonzeros are stored lane-fastest (element k of lane l at k*64 + l), so each
// load instruction of a wave reads 64 consecutive entries; bit k of lane l's
// flag word marks a row start.  A workgroup takes a unit of consecutive
// tiles of one partition in chunks of four, one tile per wave.  Per chunk
// each lane forms running sums of its sigma products that restart at its
// start bits; a block-wide segmented scan (DPP steps across each wave, the
// four wave totals through LDS) joins the sums that cross lanes and tiles,
// and the sum still open at the chunk's end carries to the next chunk in
// registers.  So y_offset and scansum_offset of CSR5's descriptor are not
// stored: the row of a segment is the tile pointer plus a prefix count of
// start bits (a binary search of the row pointer in tiles that cover empty
// rows).  Row sums leave through LDS in one coalesced pass per chunk.  The
// unit's x window is staged in LDS once per unit.  Pieces of rows cut by the
// unit or partition bounds go to the unit's two carry slots with their
// first/final tags (k_slot_fix adds them).

// upper_bound over the partition's clamped local row pointer, minus 1: the
// local row holding local position pos.
__device__ inline int tile_row_search(const int32_t *rowptr, int row_lo, int s, int nz, int lo, int hi, int pos) {
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    int v = rowptr[row_lo + mid] - s;
    v = v < 0 ? 0 : (v > nz ? nz : v);
    if (v <= pos)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - 1;
}

template <int CTRL, int ROWS, int BANKS>
__device__ inline void scan_step_c(int &cn) {
  cn += __builtin_amdgcn_update_dpp(0, cn, CTRL, ROWS, BANKS, false);
}

template <typename V, int MAXS, bool EXACT>
__global__ __launch_bounds__(kBlock) void k_tile(KArgs<V> a) {
  constexpr int NW = kBlock / kWave;
  __shared__ __attribute__((aligned(16))) double s_x[kSegStage + 2];
  __shared__ double s_wv[2][NW];  // per wave, double-buffered: scan value, closed flag, start count, dirty
  __shared__ int s_ws[2][NW];
  __shared__ int s_wc[2][NW];
  __shared__ int s_wd[2][NW];
  __shared__ double s_out[kTileOut];  // row sums closed in the current chunk, by row
  __shared__ uint8_t s_ok[kTileOut];
  __shared__ int s_unit[4];  // the unit's first row and start bit; the row and start bit after it
  __shared__ int s_kb[2];    // the chunk's first row (row sums gathered from it on)
  __shared__ int s_slot;
  const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
  // EXACT: the tile width is MAXS, so the element loops need no bounds
  const int sigma = EXACT ? MAXS : a.sigma;
  for (int t = tid; t < kTileOut; t += kBlock) s_ok[t] = 0;
  int par = 0;
  int64_t iter = 0;
  int64_t u = next_unit(a.queue, iter++, a.unit_count, &s_slot);
  int dv = desc_prefetch(a.tunits + a.unit_begin, u, a.unit_count, lane);
  while (u < a.unit_count) {
    const int dcur = dv;
    u = next_unit(a.queue, iter++, a.unit_count, &s_slot);
    dv = desc_prefetch(a.tunits + a.unit_begin, u, a.unit_count, lane);
    const int gt0 = desc_field(dcur, 0), nt = desc_field(dcur, 1), lt0 = desc_field(dcur, 2),
              tp0 = desc_field(dcur, 3), nz = desc_field(dcur, 4), row_lo = desc_field(dcur, 5),
              s0 = desc_field(dcur, 6), gu = desc_field(dcur, 7), cmin = desc_field(dcur, 8),
              cwin = desc_field(dcur, 9), tfl = desc_field(dcur, 10), last_lr = desc_field(dcur, 11);
    const bool split_head = tfl & kTuSplitHead, split_tail = tfl & kTuSplitTail;
    const bool stg = cwin > 0;
    int lead = 0;
    __syncthreads();  // the previous unit is done with s_x and s_unit
    if (stg) lead = stage_window<kBlock>(a.x, cmin, cwin, s_x, tid);
    if (tid == 0) {
      // the row open at the unit's start and the one open at its end
      const uint32_t tpa = a.tile_ptr[tp0], tpb = a.tile_ptr[tp0 + nt];
      s_unit[0] = (int)(tpa & 0x7FFFFFFFu);
      s_unit[1] = (int)(a.tile_flags[(int64_t)gt0 * kOmega] & 1u);
      s_unit[2] = (int)(tpb & 0x7FFFFFFFu);
      s_unit[3] = (int)(a.tile_flags[(int64_t)(gt0 + nt) * kOmega] & 1u);
    }
    double carry = 0.0;  // the sum open at the end of the previous chunk
    for (int c = 0; 4 * c < nt; c++, par ^= 1) {
      const int ti = 4 * c + wid;  // this wave's tile within the unit
      const bool tvalid = ti < nt;
      const int64_t gt = (int64_t)gt0 + ti;
      const int lt = lt0 + ti;
      uint32_t fl = 0, fn = 0, tpw = 0, tpn = 0;
      int32_t ci[MAXS];
      V vv[MAXS];
#pragma unroll
      for (int k = 0; k < MAXS; k++) {
        ci[k] = 0;
        vv[k] = (V)0;
      }
      if (tvalid) {
        __builtin_amdgcn_s_setprio(2);
        const int64_t base = gt * kOmega * sigma + lane;
#pragma unroll
        for (int k = 0; k < MAXS; k++) {
          if (k < sigma) {
            vv[k] = __builtin_nontemporal_load(a.val + base + (int64_t)k * kOmega);
            ci[k] = __builtin_nontemporal_load(a.col + base + (int64_t)k * kOmega);
          }
        }
        fl = __builtin_nontemporal_load(a.tile_flags + gt * kOmega + lane);
        fn = __builtin_nontemporal_load(a.tile_flags + gt * kOmega + lane + 1);  // padded by one tile
        tpw = a.tile_ptr[tp0 + ti];
        tpn = a.tile_ptr[tp0 + ti + 1];  // the tile's last row (a dirty tile's search range)
        __builtin_amdgcn_s_setprio(0);
      }
      if (c == 0) {
        if (stg) stage_wait();
        __syncthreads();  // the window and the unit's facts are in LDS
      }
      const int lrs = s_unit[0], fs = s_unit[1];
      // products of this lane's valid elements
      const int lpos = lt * kOmega * sigma + lane * sigma;  // local position of the lane's element 0
      int nv = nz - lpos;
      nv = !tvalid || nv < 0 ? 0 : (nv > sigma ? sigma : nv);
      double p[MAXS];
      if (stg) {
#pragma unroll
        for (int k = 0; k < MAXS; k++) {
          const bool ok = k < nv;
          const double t = prod(vv[k], s_x[lead + (ok ? ci[k] - cmin : 0)]);
          p[k] = ok ? t : 0.0;
        }
      } else {
#pragma unroll
        for (int k = 0; k < MAXS; k++) {
          const bool ok = k < nv;
          const double t = prod(vv[k], a.x[ok ? ci[k] : 0]);
          p[k] = ok ? t : 0.0;
        }
      }
      const uint32_t vmask = nv >= 32 ? 0xFFFFFFFFu : (1u << nv) - 1u;  // this lane's valid elements
      const uint32_t fb = fl & vmask;                                    // their start bits
      const int tclosed = lpos + nv >= nz ? 1 : (int)(fn & 1u);     // the lane's last element ends a row
      // running sums that restart at every start bit
      const uint32_t inner = fb & ~1u;
      const int f1 = inner ? __builtin_ctz(inner) : nv;
#pragma unroll
      for (int k = 1; k < MAXS; k++)
        if (!((fb >> k) & 1u)) p[k] = p[k - 1] + p[k];
      double last = p[0];
#pragma unroll
      for (int k = 1; k < MAXS; k++)
        if (k < nv) last = p[k];
      const int has_close = nv > 0 && (inner != 0 || tclosed);
      const double run = nv > 0 && !tclosed ? last : 0.0;
      const int cnt = __builtin_popcount(fb);
      // segmented scan of the running sums across the workgroup, and a wave
      // scan of start-bit counts (rows of this tile)
      double v = run;
      int sg = has_close, cn = cnt;
      scan_step<0x111, 0xf, 0xf>(v, sg, cn);  // row_shr:1
      scan_step<0x112, 0xf, 0xf>(v, sg, cn);  // row_shr:2
      scan_step<0x114, 0xf, 0xe>(v, sg, cn);  // row_shr:4
      scan_step<0x118, 0xf, 0xc>(v, sg, cn);  // row_shr:8
      scan_step<0x142, 0xa, 0xf>(v, sg, cn);  // row_bcast:15 into rows 1, 3
      scan_step<0x143, 0xc, 0xf>(v, sg, cn);  // row_bcast:31 into rows 2, 3
      const bool dirty = tvalid && (tpw >> 31);
      const int R0 = (int)(tpw & 0x7FFFFFFFu);
      const int f00 = (int)(__builtin_amdgcn_readlane(fl, 0) & 1u);
      if (lane == kWave - 1) {
        s_wv[par][wid] = v;
        s_ws[par][wid] = sg;
        s_wc[par][wid] = cn;
        s_wd[par][wid] = dirty ? 1 : 0;
      }
      if (tid == 0) s_kb[par] = R0;
      __syncthreads();
      double prev = carry, tot = carry;
      int totc = 0, anyd = 0;
#pragma unroll
      for (int w = 0; w < NW; w++) {
        const double wv = s_wv[par][w];
        const int ws = s_ws[par][w];
        if (w < wid) prev = wv + (ws ? 0.0 : prev);
        tot = wv + (ws ? 0.0 : tot);
        totc += s_wc[par][w];
        anyd |= s_wd[par][w];
      }
      const double c_incl = v + (sg ? 0.0 : prev);
      double carry_in = __shfl_up(c_incl, 1, kWave);
      if (lane == 0) carry_in = prev;
      const int excl = cn - cnt;  // start bits in earlier lanes of this tile
      // rows closed in this chunk lie in [kb, kb + totc] unless a tile
      // covers empty rows; they go through LDS when that span fits
      const int kb = s_kb[par];
      const bool via_lds = !anyd && totc + 1 <= kTileOut;
      const uint32_t ends = nv > 0 ? ((fb >> 1) | ((uint32_t)tclosed << (nv - 1))) & vmask : 0u;
      // the local rows this lane can close: [lo_lr, lo_lr + its start bits]
      // (clean tiles); none of them a carry piece: every sum goes to LDS
      const int lo_lr = R0 + excl - f00;
      const int hi_lr = lo_lr + cnt;
      const bool plain = via_lds && !dirty && !(!fs && lrs >= lo_lr && lrs <= hi_lr) && !(split_head && lo_lr <= 0) &&
                         !(split_tail && hi_lr >= last_lr);
      if (plain) {
#pragma unroll
        for (int k = 0; k < MAXS; k++) {
          if ((ends >> k) & 1u) {
            const int t = lo_lr + __builtin_popcount(fb & ((2u << k) - 1u)) - kb;
            s_out[t] = k < f1 ? p[k] + carry_in : p[k];
            s_ok[t] = 1;
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < MAXS; k++) {
          if ((ends >> k) & 1u) {
            const double val = k < f1 ? p[k] + carry_in : p[k];
            const int lr = dirty ? tile_row_search(a.rowptr, row_lo, s0, nz, R0, (int)(tpn & 0x7FFFFFFFu) + 1, lpos + k)
                                 : R0 + excl + __builtin_popcount(fb & ((2u << k) - 1u)) - f00;
            const int r = row_lo + lr;
            // the pieces cut by the unit's start or the partition's bounds go
            // to carry slots, the rows complete here to y
            if (lr == lrs && !fs) {
              const bool fin = !(split_tail && lr == last_lr);
              a.crow[2 * (int64_t)gu] = (int32_t)((uint32_t)r | (fin ? kSlotFinal : 0u));
              a.cval[2 * (int64_t)gu] = val;
            } else if (split_head && lr == 0) {
              const bool fin = !(split_tail && lr == last_lr);
              a.crow[2 * (int64_t)gu] = (int32_t)((uint32_t)r | (fin ? kSlotFinal : 0u));
              a.cval[2 * (int64_t)gu] = val;
            } else if (split_tail && lr == last_lr) {
              a.crow[2 * (int64_t)gu + 1] = r | kSlotFirst;
              a.cval[2 * (int64_t)gu + 1] = val;
            } else if (via_lds) {
              s_out[lr - kb] = val;
              s_ok[lr - kb] = 1;
            } else {
              a.y[r] = y_update(a.alpha, a.beta, val, a.y, r);
            }
          }
        }
      }
      if (via_lds) {
        // s_out is written only after the next chunk's first barrier, which
        // every thread reaches after this pass: one buffer suffices
        __syncthreads();
        for (int t = tid; t <= totc; t += kBlock) {
          if (s_ok[t]) {
            const int r = row_lo + kb + t;
            a.y[r] = y_update(a.alpha, a.beta, s_out[t], a.y, r);
            s_ok[t] = 0;
          }
        }
      }
      carry = tot;
    }
    // the sum open at the unit's end: the row holding the next tile's first
    // element continues, unless that element starts a row or the partition
    // ends here
    if (tid == 0) {
      const int64_t uend = (int64_t)(lt0 + nt) * kOmega * sigma;
      if (uend < nz && !s_unit[3]) {
        const int lr = s_unit[2];
        const bool began = !(lr == s_unit[0] && !s_unit[1]);  // else the whole unit lies in the row
        const bool first = began && !(split_head && lr == 0);
        const int64_t slot = 2 * (int64_t)gu + (began ? 1 : 0);
        a.crow[slot] = (row_lo + lr) | (first ? kSlotFirst : 0);
        a.cval[slot] = carry;
      }
    }
  }
}

// ---- x windows of units (plan time) -----------------------------------------------
__global__ __launch_bounds__(kBlock) void k_unit_window(const int32_t *__restrict__ col, int32_t *__restrict__ f,
                                                        int stride, int off, int64_t nunits) {
  __shared__ int s_lo[kBlock / kWave], s_hi[kBlock / kWave];
  const int64_t u = blockIdx.x;
  if (u >= nunits) return;
  const int lo = f[u * stride + off], hi = f[u * stride + off + 1];
  int mn = INT32_MAX, mx = -1;
  for (int k = lo + (int)threadIdx.x; k < hi; k += kBlock) {
    const int c = col[k];
    mn = c < mn ? c : mn;
    mx = c > mx ? c : mx;
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const int a2 = __shfl_xor(mn, o, kWave), b2 = __shfl_xor(mx, o, kWave);
    mn = a2 < mn ? a2 : mn;
    mx = b2 > mx ? b2 : mx;
  }
  if (threadIdx.x % kWave == 0) {
    s_lo[threadIdx.x / kWave] = mn;
    s_hi[threadIdx.x / kWave] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / kWave; w++) {
      mn = s_lo[w] < mn ? s_lo[w] : mn;
      mx = s_hi[w] > mx ? s_hi[w] : mx;
    }
    const int64_t win = mx >= mn ? (int64_t)mx - mn + 1 : 0;
    f[u * stride + off] = win > 0 && win <= kSegStage ? mn : 0;
    f[u * stride + off + 1] = win > 0 && win <= kSegStage ? (int)win : 0;
  }
}

// ---- launchers --------------------------------------------------------------------

static inline int nblk(int64_t n, int b = kBlock) { return (int)((n + b - 1) / b); }

hipError_t launch_partition(const int32_t *rowptr, int m, int64_t nnz, int P, int kind, Part *parts,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_partition, dim3(nblk(P)), dim3(kBlock), 0, st, rowptr, m, nnz, P, kind, parts);
  return hipGetLastError();
}

hipError_t launch_merge_coords(const int32_t *rowptr, const Part *parts, int nparts, int64_t total,
                               int tw, Coord *coords, hipStream_t st) {
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_merge_coords, dim3(nblk(total)), dim3(kBlock), 0, st, rowptr, parts, nparts,
                     total, tw, coords);
  return hipGetLastError();
}

hipError_t launch_tile_convert(const int32_t *rowptr, const int32_t *col, const double *val,
                               const Part *hparts, const Part *dparts, int nparts, int sigma,
                               uint32_t *tile_ptr, uint32_t *flags, int32_t *colt, double *valt,
                               float *valt32, hipStream_t st) {
  const Part &last = hparts[nparts - 1];
  const int64_t ntiles = last.tile_off + last.ntiles;
  const int64_t ntp = last.tp_off + last.ntiles + 1;
  hipLaunchKernelGGL(k_tile_ptr, dim3(nblk(ntp)), dim3(kBlock), 0, st, rowptr, dparts, nparts, ntp,
                     sigma, tile_ptr);
  if (ntiles > 0) {
    hipLaunchKernelGGL(k_tile_dirty, dim3(nblk(ntiles)), dim3(kBlock), 0, st, rowptr, dparts, nparts,
                       ntiles, tile_ptr);
    hipLaunchKernelGGL(k_tile_flags, dim3(nblk(ntiles * kOmega)), dim3(kBlock), 0, st, rowptr, dparts,
                       nparts, ntiles * kOmega, sigma, flags);
    const int64_t tot = ntiles * kOmega * sigma;
    if (valt32)
      hipLaunchKernelGGL(k_tile_transpose<float>, dim3(nblk(tot)), dim3(kBlock), 0, st, dparts, nparts,
                         tot, sigma, col, val, colt, valt32);
    else
      hipLaunchKernelGGL(k_tile_transpose<double>, dim3(nblk(tot)), dim3(kBlock), 0, st, dparts, nparts,
                         tot, sigma, col, val, colt, valt);
  }
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n)), dim3(kBlock), 0, st, p, v, n);
  return hipGetLastError();
}

hipError_t launch_to_f32(const double *v, float *o, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_f32, dim3(nblk(n)), dim3(kBlock), 0, st, v, o, n);
  return hipGetLastError();
}

hipError_t launch_empty_rows(const int32_t *rows, int64_t n, double beta, double *y, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_empty_rows, dim3(nblk(n)), dim3(kBlock), 0, st, rows, n, beta, y);
  return hipGetLastError();
}

hipError_t launch_find_empty(const int32_t *rowptr, int m, int32_t *flag, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_find_empty, dim3(nblk(m)), dim3(kBlock), 0, st, rowptr, m, flag);
  return hipGetLastError();
}

hipError_t launch_long_reduce(const double *partial, int nlong, const int32_t *lsp, const int32_t *lseg,
                              const int32_t *rows, double alpha, double beta, bool assign, double *y,
                              hipStream_t st) {
  if (nlong <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_long_reduce, dim3(nlong), dim3(kBlock), 0, st, partial, nlong, lsp, lseg, rows, alpha, beta,
                     assign ? 1 : 0, y);
  return hipGetLastError();
}

hipError_t launch_reset(int *q, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_reset, dim3(nblk(n)), dim3(kBlock), 0, st, q, n);
  return hipGetLastError();
}

// Workgroups that fit resident on the whole device for kernel `k`
// (occupancy x CUs), rounded down to a multiple of kXcd.
static int resident_cap(const void *k) {
  int nb = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, kBlock, 0) != hipSuccess || nb < 1) nb = 1;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  int cap = nb * cus / kXcd * kXcd;
  return cap < kXcd ? kXcd : cap;
}

// Grid for a launch: at most what fits resident, at least one workgroup per
// XCD, always a multiple of kXcd (the unit walk divides by it).
static int launch_grid(int want, int cap) {
  int g = want < cap ? want : cap;
  g = g / kXcd * kXcd;
  return g < kXcd ? kXcd : g;
}

#define RESIDENT(kern) ([] { static const int c = resident_cap((const void *)(kern)); return c; }())

template <typename V>
static hipError_t vec_dispatch(const KArgs<V> &a, int lanes, int grid, hipStream_t st) {
#define SPX_VEC(L)                                                                                   \
  hipLaunchKernelGGL((k_vector<V, L>), dim3(launch_grid(grid, RESIDENT((k_vector<V, L>)))), dim3(kBlock), \
                     0, st, a)
  switch (lanes) {
    case 2: SPX_VEC(2); break;
    case 4: SPX_VEC(4); break;
    case 8: SPX_VEC(8); break;
    case 16: SPX_VEC(16); break;
    case 32: SPX_VEC(32); break;
    case 64: SPX_VEC(64); break;
    default: return hipErrorInvalidValue;
  }
#undef SPX_VEC
  return hipGetLastError();
}

template <typename V>
static hipError_t merge_dispatch(const KArgs<V> &a, int items, int grid, hipStream_t st) {
#define SPX_MERGE(I) \
  hipLaunchKernelGGL((k_merge<V, I>), dim3(launch_grid(grid, RESIDENT((k_merge<V, I>)))), dim3(kBlock), 0, st, a)
  switch (items) {
    case 2: SPX_MERGE(2); break;
    case 4: SPX_MERGE(4); break;
    case 6: SPX_MERGE(6); break;
    case 8: SPX_MERGE(8); break;
    case 12: SPX_MERGE(12); break;
    case 16: SPX_MERGE(16); break;
    default: return hipErrorInvalidValue;
  }
#undef SPX_MERGE
  return hipGetLastError();
}

hipError_t launch_vector_f64(const KArgs<double> &a, int lanes, int grid, hipStream_t st) {
  return vec_dispatch(a, lanes, grid, st);
}
hipError_t launch_vector_f32(const KArgs<float> &a, int lanes, int grid, hipStream_t st) {
  return vec_dispatch(a, lanes, grid, st);
}
hipError_t launch_merge_f64(const KArgs<double> &a, int items, int grid, hipStream_t st) {
  return merge_dispatch(a, items, grid, st);
}
hipError_t launch_merge_f32(const KArgs<float> &a, int items, int grid, hipStream_t st) {
  return merge_dispatch(a, items, grid, st);
}
template <typename V>
static hipError_t tile_dispatch(const KArgs<V> &a, int grid, hipStream_t st) {
#define SPX_TILE(S, E) \
  hipLaunchKernelGGL((k_tile<V, S, E>), dim3(launch_grid(grid, RESIDENT((k_tile<V, S, E>)))), dim3(kBlock), 0, st, a)
  // the widths the automatic choice gives short rows run with exact loops
  switch (a.sigma) {
    case 4: SPX_TILE(4, true); break;
    case 5: SPX_TILE(5, true); break;
    case 6: SPX_TILE(6, true); break;
    case 7: SPX_TILE(7, true); break;
    case 8: SPX_TILE(8, true); break;
    case 16: SPX_TILE(16, true); break;
    case 32: SPX_TILE(32, true); break;
    default:
      if (a.sigma <= 8)
        SPX_TILE(8, false);
      else if (a.sigma <= 16)
        SPX_TILE(16, false);
      else
        SPX_TILE(32, false);
  }
#undef SPX_TILE
  return hipGetLastError();
}
hipError_t launch_tile_f64(const KArgs<double> &a, int grid, hipStream_t st) {
  return tile_dispatch(a, grid, st);
}
hipError_t launch_tile_f32(const KArgs<float> &a, int grid, hipStream_t st) {
  return tile_dispatch(a, grid, st);
}

hipError_t launch_finish(const double *partial, int nlong, const int32_t *lsp, const int32_t *lseg,
                         const int32_t *rows, bool assign, const int32_t *crow, const double *cval, int64_t nslots,
                         double alpha, double beta, double *y, hipStream_t st) {
  const int64_t g = nlong + (nslots + kBlock - 1) / kBlock;
  if (g <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_finish, dim3((unsigned)g), dim3(kBlock), 0, st, partial, nlong, lsp, lseg, rows, assign ? 1 : 0,
                     crow, cval, nslots, alpha, beta, y);
  return hipGetLastError();
}

hipError_t launch_slot_fix(const int32_t *crow, const double *cval, int64_t nslots, double alpha, double beta,
                           double *y, hipStream_t st) {
  if (nslots <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_slot_fix, dim3(nblk(nslots)), dim3(kBlock), 0, st, crow, cval, nslots, alpha, beta, y);
  return hipGetLastError();
}

hipError_t launch_unit_window(const int32_t *col, int32_t *f, int stride, int off, int64_t nunits, hipStream_t st) {
  if (nunits <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_unit_window, dim3((unsigned)nunits), dim3(kBlock), 0, st, col, f, stride, off, nunits);
  return hipGetLastError();
}

hipError_t launch_merge_units(const Coord *coords, int64_t units, MUnit *out, hipStream_t st) {
  if (units <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_merge_units, dim3(nblk(units)), dim3(kBlock), 0, st, coords, units, out);
  return hipGetLastError();
}

}  // namespace spx
