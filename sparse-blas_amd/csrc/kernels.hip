// kernels.hip -- hand-written gfx950 (CDNA4, wave64) kernels for
// y = alpha*A*x + beta*y on a CSR matrix, plus the device-side partitioner
// and the CSR5-style tile conversion.
//
// Work is cut into partitions (nnz-balanced or row blocks) and each partition
// into workgroup "units".  A row whose nonzeros fall in more than one unit is
// never written by a unit: every unit writes its piece of such a row into a
// carry slot, and k_fixup sums each row's carries in slot order (nonzero
// order) and applies alpha and the beta term once.  Rows a unit holds whole
// are written directly.  Summation order is fixed, so results are bitwise
// reproducible; no atomics touch y.
//
// beta == 0 never reads y (BLAS convention), so y may hold NaN on input.
#include "spx_device.h"

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

__global__ void k_split_x(const double *__restrict__ x, XPair *__restrict__ xhl, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double v = x[i];
    float h = (float)v;
    xhl[i] = XPair{h, (float)(v - (double)h)};
  }
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

// Sum each carried row's slots in slot order.  Which slots hold which row is
// fixed by the sparsity structure alone, so the plan records the runs once
// (row, first slot, end slot) and one thread per run does the sum.  Slots
// inside a run that belong to no row (-1) are skipped.
__global__ void k_fixup(const Run *__restrict__ runs, int64_t nruns, const int32_t *__restrict__ crow,
                        const double *__restrict__ cval, double alpha, double beta, double *__restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nruns) return;
  const Run R = runs[i];
  double s = 0.0;
  for (int64_t k = R.beg; k < R.end; k++)
    if (crow[k] == R.row) s += cval[k];
  y[R.row] = y_update(alpha, beta, s, y, R.row);
}

// ---- long rows: fold the per-column-block partial sums -----------------------
// The segmented engine sweeps the long rows by column blocks and leaves one
// partial sum per (column block, row) slice; lseg[lsp[r] .. lsp[r+1]) lists
// row r's slices in column order.  One workgroup per long row adds them in a
// fixed tree order, then y[row] = alpha * sum + beta * y[row] (`assign`), or
// y[row] += alpha * sum when the regular pass has already written beta * y0
// for the row (it is empty there).
__global__ __launch_bounds__(kBlock) void k_long_reduce(const double *__restrict__ partial, int nlong,
                                                        const int32_t *__restrict__ lsp,
                                                        const int32_t *__restrict__ lseg,
                                                        const int32_t *__restrict__ rows, double alpha,
                                                        double beta, int assign, double *__restrict__ y) {
  // Up to U slices per thread per round: all their slot indices are loaded
  // first, then all the partials, so a row of up to U*kBlock slices costs two
  // dependent memory round trips.  A thread adds its slices in slot order.
  constexpr int U = 16;
  __shared__ double s_w[kBlock / kWave];
  const int r = blockIdx.x, tid = threadIdx.x;
  if (r >= nlong) return;
  const int j0 = lsp[r], j1 = lsp[r + 1];
  double s = 0.0;
  for (int j = j0 + tid; j < j1; j += U * kBlock) {
    int q[U];
#pragma unroll
    for (int k = 0; k < U; k++) q[k] = j + k * kBlock < j1 ? lseg[j + k * kBlock] : -1;
    double v[U];
#pragma unroll
    for (int k = 0; k < U; k++) v[k] = q[k] >= 0 ? partial[q[k]] : 0.0;
#pragma unroll
    for (int k = 0; k < U; k++) s += v[k];
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
  if (tid % kWave == 0) s_w[tid / kWave] = s;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / kWave; w++) t += s_w[w];
    const int row = rows[r];
    if (!assign)
      y[row] += alpha * t;
    else
      y[row] = beta == 0.0 ? alpha * t : alpha * t + beta * y[row];
  }
}

// ---- kernel 1: CSR-vector, L lanes per row --------------------------------------
// A group of L lanes walks one row with stride L, then reduces with
// cross-lane shuffles (Bell & Garland's CSR-vector, wave64 edition).  The two
// rows split at the partition boundary go to the partition's carry slots.

template <typename V, typename X, int L>
__global__ __launch_bounds__(kBlock) void k_vector(KArgs<V, X> a) {
  constexpr int G = kBlock / L;
  __shared__ int s_slot;
  const int g = threadIdx.x / L, l = threadIdx.x % L;
  for (int64_t iter = 0;; iter++) {
    int64_t u = next_unit(a.queue, iter, a.unit_count, &s_slot);
    if (u >= a.unit_count) break;
    const Unit U = a.units[a.unit_begin + u];
    const Part q = a.parts[U.part];  // by value: y stores cannot alias it
    const int64_t r = (int64_t)q.row_lo + (int64_t)U.local * G + g;
    if (r < q.row_hi) {
      int64_t b = a.rowptr[r];
      if (b < q.s) b = q.s;
      int64_t e = a.rowptr[r + 1];
      if (e > q.e + 1) e = q.e + 1;
      double sum = 0.0, sum2 = 0.0;
      int64_t j = b + l;
      for (; j + L < e; j += 2 * L) {
        sum += prod(a.val[j], a.x, a.col[j]);
        sum2 += prod(a.val[j + L], a.x, a.col[j + L]);
      }
      if (j < e) sum += prod(a.val[j], a.x, a.col[j]);
      sum += sum2;
#pragma unroll
      for (int off = L / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, L);
      if (l == 0) {
        if (r == q.r0 && q.sflag)
          put_carry(a.crow, a.cval, q.carry_off, (int)r, sum);
        else if (r == q.r1 && q.eflag)
          put_carry(a.crow, a.cval, q.cslot1, (int)r, sum);
        else
          a.y[r] = y_update(a.alpha, a.beta, sum, a.y, (int)r);
      }
    }
    if (a.queue) __syncthreads();
  }
}

// ---- kernel 2: merge path ------------------------------------------------------
// Merrill & Garland (SC'16): the merge of row ends and nonzero indices is cut
// into equal pieces of kBlock*IPT items, one per workgroup; unit boundary
// coordinates come from k_merge_coords at plan time.  A workgroup stages its
// row ends and its products val*x[col] in LDS with coalesced loads, each
// thread merges IPT consecutive items, and a wave-shuffle segmented scan
// carries partial row sums across threads.

template <typename V, typename X, int IPT>
__global__ __launch_bounds__(kBlock) void k_merge(KArgs<V, X> a) {
  constexpr int TW = kBlock * IPT;
  __shared__ double s_prod[TW];
  __shared__ int32_t s_re[TW + 1];
  __shared__ double s_wv[kBlock / kWave];
  __shared__ int s_ws[kBlock / kWave];
  __shared__ int s_slot;
  const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;

  for (int64_t iter = 0;; iter++) {
    int64_t u = next_unit(a.queue, iter, a.unit_count, &s_slot);
    if (u >= a.unit_count) break;
    const Unit U = a.units[a.unit_begin + u];
    const Part q = a.parts[U.part];  // by value: y stores cannot alias it
    const Coord c0 = a.coords[q.coord_off + U.local];
    const Coord c1 = a.coords[q.coord_off + U.local + 1];
    const int i0 = c0.row, j0 = c0.nz;
    const int nrows = c1.row - c0.row, nnzs = c1.nz - c0.nz;

    // stage row ends and products (coalesced; gathers of x issued together)
    for (int k = tid; k < nrows; k += kBlock) s_re[k] = a.rowptr[i0 + k + 1];
    {
      int cidx[IPT];
      V vv[IPT];
#pragma unroll
      for (int r = 0; r < IPT; r++) {
        int k = tid + r * kBlock;
        cidx[r] = k < nnzs ? a.col[j0 + k] : 0;
        vv[r] = k < nnzs ? a.val[j0 + k] : (V)0;
      }
#pragma unroll
      for (int r = 0; r < IPT; r++) {
        int k = tid + r * kBlock;
        if (k < nnzs) s_prod[k] = prod(vv[r], a.x, cidx[r]);
      }
    }
    __syncthreads();

    // this thread's merge items [d, dend)
    const int total = nrows + nnzs;
    int d = tid * IPT;
    if (d > total) d = total;
    int dend = d + IPT;
    if (dend > total) dend = total;
    int lo = d - nnzs > 0 ? d - nnzs : 0, hi = d < nrows ? d : nrows;
    while (lo < hi) {
      int piv = (lo + hi) >> 1;
      if (s_re[piv] <= j0 + (d - piv - 1))
        lo = piv + 1;
      else
        hi = piv;
    }
    int it = lo, jt = d - lo;
    const int first_row = it;
    double run = 0.0, first_val = 0.0;
    int has_end = 0;
    for (int k = d; k < dend; k++) {
      if (it < nrows && (jt >= nnzs || s_re[it] <= j0 + jt)) {
        if (!has_end) {
          first_val = run;
          has_end = 1;
        } else {
          a.y[i0 + it] = y_update(a.alpha, a.beta, run, a.y, i0 + it);
        }
        run = 0.0;
        it++;
      } else {
        run += s_prod[jt];
        jt++;
      }
    }

    // segmented inclusive scan c(t) = run(t) + (has_end(t) ? 0 : c(t-1))
    double v = run;
    int sg = has_end;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      double vu = __shfl_up(v, off, kWave);
      int su = __shfl_up(sg, off, kWave);
      if (lane >= off && !sg) {
        v += vu;
        sg |= su;
      }
    }
    if (lane == kWave - 1) {
      s_wv[wid] = v;
      s_ws[wid] = sg;
    }
    __syncthreads();
    double prev = 0.0;  // inclusive value at the end of the previous wave
    for (int w = 0; w < wid; w++) prev = s_wv[w] + (s_ws[w] ? 0.0 : prev);
    const double c_incl = v + (sg ? 0.0 : prev);
    double carry_in = __shfl_up(c_incl, 1, kWave);
    if (lane == 0) carry_in = prev;

    const int64_t slot0 = q.carry_off + 1 + 2 * (int64_t)U.local;
    if (has_end) {
      const double tot = first_val + carry_in;
      const int r = i0 + first_row;
      if (first_row == 0 && a.rowptr[i0] < j0)
        put_carry(a.crow, a.cval, slot0, r, tot);  // row began in an earlier unit
      else
        a.y[r] = y_update(a.alpha, a.beta, tot, a.y, r);
    }
    if (tid == kBlock - 1) {
      // tail: nonzeros after the last row end belong to the row still open
      const int tail_from = nrows > 0 ? s_re[nrows - 1] : j0;
      const bool has_tail = c1.nz > tail_from;
      const bool head_split = nrows > 0 && a.rowptr[i0] < j0;
      if (nrows == 0) {
        if (has_tail)
          put_carry(a.crow, a.cval, slot0, c1.row, c_incl);
        else
          put_carry(a.crow, a.cval, slot0, -1, 0.0);
        put_carry(a.crow, a.cval, slot0 + 1, -1, 0.0);
      } else {
        if (!head_split) put_carry(a.crow, a.cval, slot0, -1, 0.0);
        put_carry(a.crow, a.cval, slot0 + 1, has_tail ? c1.row : -1, has_tail ? c_incl : 0.0);
      }
    }
    __syncthreads();
  }
}

// ---- kernel 3: CSR5-style tiles -------------------------------------------------
// One wave64 per tile of 64 lanes x sigma nonzeros stored lane-fastest, so
// every load instruction reads 64 consecutive entries.  Each lane walks its
// sigma elements; bit k of the lane's flag word marks a row start.  Row ids
// come from the tile pointer plus a wave prefix count of the flags (or, in
// tiles flagged as covering empty rows, a binary search of the row pointer).
// The open row at a lane's end is completed by a backward segmented scan of
// the following lanes' leading sums (Liu & Vinter, ICS'15, recast for wave64
// with ballots and shuffles instead of stored y/scansum offsets).

template <typename V, typename X>
__device__ inline void tile_emit(const KArgs<V, X> &a, const Part &q, int64_t lrow, double v) {
  const int r = (int)(q.row_lo + lrow);
  if (r == q.r0 && q.sflag)
    put_carry(a.crow, a.cval, q.carry_off, r, v);
  else if (r == q.r1 && q.eflag)
    put_carry(a.crow, a.cval, q.cslot1, r, v);
  else
    a.y[r] = y_update(a.alpha, a.beta, v, a.y, r);
}

template <typename V, typename X, int MAXS>
__global__ __launch_bounds__(kBlock) void k_tile(KArgs<V, X> a) {
  __shared__ int s_slot;
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  const int sigma = a.sigma;
  for (int64_t iter = 0;; iter++) {
    int64_t u = next_unit(a.queue, iter, a.unit_count, &s_slot);
    if (u >= a.unit_count) break;
    const Unit U = a.units[a.unit_begin + u];
    const Part q = a.parts[U.part];  // by value: y stores cannot alias it
    const int64_t t = (int64_t)U.local * (kBlock / kWave) + wid;
    if (t < q.ntiles) {
      const int64_t nz = q.e - q.s + 1;
      const uint32_t tpw = a.tile_ptr[q.tp_off + t];
      const bool dirty = tpw >> 31;
      const int64_t R0 = tpw & 0x7FFFFFFFu;
      const int64_t R1 = a.tile_ptr[q.tp_off + t + 1] & 0x7FFFFFFFu;
      const int64_t tstart = t * kOmega * sigma;
      int64_t tend = tstart + (int64_t)kOmega * sigma;
      if (tend > nz) tend = nz;
      const uint32_t fl = a.tile_flags[(q.tile_off + t) * kOmega + lane];
      // does a row start exactly where this tile ends?  (bit 0 of the next
      // tile's lane 0; a row open at the partition end does not continue)
      const bool next_starts = tend >= nz || (a.tile_flags[(q.tile_off + t + 1) * kOmega] & 1u);
      // exclusive prefix count of row starts over lanes
      int excl = __popc(fl);
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        int o = __shfl_up(excl, off, kWave);
        if (lane >= off) excl += o;
      }
      excl -= __popc(fl);
      const int f00 = __shfl(fl, 0, kWave) & 1;
      int64_t cur = R0 + excl + (f00 ? 0 : 1) - 1;  // row open at the lane start
      const int64_t base = (q.tile_off + t) * kOmega * sigma;
      double sum = 0.0, head = 0.0;
      int seen = 0;
      const int64_t lpos = tstart + (int64_t)lane * sigma;
      // all column/value loads first, then all x gathers, then the
      // flag-driven segmented sum out of registers (the tile storage is
      // padded to whole tiles, so loads need no bound; products past nz do)
      double p[MAXS];
      {
        int c[MAXS];
        V v[MAXS];
#pragma unroll
        for (int k = 0; k < MAXS; k++) {
          c[k] = 0;
          v[k] = (V)0;
          if (k < sigma) {
            const int64_t idx = base + (int64_t)k * kOmega + lane;
            c[k] = a.col[idx];
            v[k] = a.val[idx];
          }
        }
#pragma unroll
        for (int k = 0; k < MAXS; k++) p[k] = (k < sigma && lpos + k < nz) ? prod(v[k], a.x, c[k]) : 0.0;
      }
#pragma unroll
      for (int k = 0; k < MAXS; k++) {
        if (k < sigma) {
          if ((fl >> k) & 1u) {
            if (seen)
              tile_emit(a, q, cur, sum);
            else
              head = sum;
            seen = 1;
            sum = 0.0;
            cur = dirty ? local_row_search(a.rowptr, q, R0, R1 + 1, lpos + k) : cur + 1;
          }
          sum += p[k];
        }
      }
      double tail = 0.0;
      if (seen)
        tail = sum;
      else
        head = sum;
      // T(l) = head(l) + (seen(l) ? 0 : T(l+1)): backward segmented scan
      double tv = head;
      int ts = seen;
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        double vd = __shfl_down(tv, off, kWave);
        int sd = __shfl_down(ts, off, kWave);
        if (lane + off < kWave && !ts) {
          tv += vd;
          ts |= sd;
        }
      }
      double tnext = __shfl_down(tv, 1, kWave);
      if (lane == kWave - 1) tnext = 0.0;
      const uint64_t mask = __ballot(seen);
      const int64_t slot0 = q.carry_off + 1 + 2 * t;
      const bool later = lane < kWave - 1 && (mask >> (lane + 1)) != 0;
      if (seen) {
        const double tot = tail + tnext;
        if (!later) {
          // the last row started in this tile: carried on if it runs past the tile
          if (!next_starts) {
            put_carry(a.crow, a.cval, slot0 + 1, (int)(q.row_lo + cur), tot);
          } else {
            tile_emit(a, q, cur, tot);
            put_carry(a.crow, a.cval, slot0 + 1, -1, 0.0);
          }
        } else {
          tile_emit(a, q, cur, tot);
        }
      }
      if (lane == 0) {
        // leading piece of a row begun in an earlier tile
        if (!f00)
          put_carry(a.crow, a.cval, slot0, (int)(q.row_lo + R0), tv);
        else
          put_carry(a.crow, a.cval, slot0, -1, 0.0);
        if (mask == 0) put_carry(a.crow, a.cval, slot0 + 1, -1, 0.0);
      }
    }
    if (a.queue) __syncthreads();
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

hipError_t launch_split_x(const double *x, XPair *xhl, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_split_x, dim3(nblk(n)), dim3(kBlock), 0, st, x, xhl, n);
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

hipError_t launch_fixup(const Run *runs, int64_t nruns, const int32_t *crow, const double *cval, double alpha,
                        double beta, double *y, hipStream_t st) {
  if (nruns <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fixup, dim3(nblk(nruns)), dim3(kBlock), 0, st, runs, nruns, crow, cval, alpha, beta, y);
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

template <typename V, typename X>
static hipError_t vec_dispatch(const KArgs<V, X> &a, int lanes, int grid, hipStream_t st) {
  switch (lanes) {
    case 1: hipLaunchKernelGGL((k_vector<V, X, 1>), dim3(launch_grid(grid, RESIDENT((k_vector<V, X, 1>)))), dim3(kBlock), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_vector<V, X, 2>), dim3(launch_grid(grid, RESIDENT((k_vector<V, X, 2>)))), dim3(kBlock), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_vector<V, X, 4>), dim3(launch_grid(grid, RESIDENT((k_vector<V, X, 4>)))), dim3(kBlock), 0, st, a); break;
    case 8: hipLaunchKernelGGL((k_vector<V, X, 8>), dim3(launch_grid(grid, RESIDENT((k_vector<V, X, 8>)))), dim3(kBlock), 0, st, a); break;
    case 16: hipLaunchKernelGGL((k_vector<V, X, 16>), dim3(launch_grid(grid, RESIDENT((k_vector<V, X, 16>)))), dim3(kBlock), 0, st, a); break;
    case 32: hipLaunchKernelGGL((k_vector<V, X, 32>), dim3(launch_grid(grid, RESIDENT((k_vector<V, X, 32>)))), dim3(kBlock), 0, st, a); break;
    case 64: hipLaunchKernelGGL((k_vector<V, X, 64>), dim3(launch_grid(grid, RESIDENT((k_vector<V, X, 64>)))), dim3(kBlock), 0, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename V, typename X>
static hipError_t merge_dispatch(const KArgs<V, X> &a, int items, int grid, hipStream_t st) {
  switch (items) {
    case 2: hipLaunchKernelGGL((k_merge<V, X, 2>), dim3(launch_grid(grid, RESIDENT((k_merge<V, X, 2>)))), dim3(kBlock), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_merge<V, X, 4>), dim3(launch_grid(grid, RESIDENT((k_merge<V, X, 4>)))), dim3(kBlock), 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_merge<V, X, 6>), dim3(launch_grid(grid, RESIDENT((k_merge<V, X, 6>)))), dim3(kBlock), 0, st, a); break;
    case 8: hipLaunchKernelGGL((k_merge<V, X, 8>), dim3(launch_grid(grid, RESIDENT((k_merge<V, X, 8>)))), dim3(kBlock), 0, st, a); break;
    case 12: hipLaunchKernelGGL((k_merge<V, X, 12>), dim3(launch_grid(grid, RESIDENT((k_merge<V, X, 12>)))), dim3(kBlock), 0, st, a); break;
    case 16: hipLaunchKernelGGL((k_merge<V, X, 16>), dim3(launch_grid(grid, RESIDENT((k_merge<V, X, 16>)))), dim3(kBlock), 0, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_vector_f64(const KArgs<double, double> &a, int lanes, int grid, hipStream_t st) {
  return vec_dispatch(a, lanes, grid, st);
}
hipError_t launch_vector_f32(const KArgs<float, XPair> &a, int lanes, int grid, hipStream_t st) {
  return vec_dispatch(a, lanes, grid, st);
}
hipError_t launch_merge_f64(const KArgs<double, double> &a, int items, int grid, hipStream_t st) {
  return merge_dispatch(a, items, grid, st);
}
hipError_t launch_merge_f32(const KArgs<float, XPair> &a, int items, int grid, hipStream_t st) {
  return merge_dispatch(a, items, grid, st);
}
template <typename V, typename X>
static hipError_t tile_dispatch(const KArgs<V, X> &a, int grid, hipStream_t st) {
  if (a.sigma <= 8)
    hipLaunchKernelGGL((k_tile<V, X, 8>), dim3(launch_grid(grid, RESIDENT((k_tile<V, X, 8>)))), dim3(kBlock), 0, st, a);
  else if (a.sigma <= 16)
    hipLaunchKernelGGL((k_tile<V, X, 16>), dim3(launch_grid(grid, RESIDENT((k_tile<V, X, 16>)))), dim3(kBlock), 0, st, a);
  else
    hipLaunchKernelGGL((k_tile<V, X, 32>), dim3(launch_grid(grid, RESIDENT((k_tile<V, X, 32>)))), dim3(kBlock), 0, st, a);
  return hipGetLastError();
}
hipError_t launch_tile_f64(const KArgs<double, double> &a, int grid, hipStream_t st) {
  return tile_dispatch(a, grid, st);
}
hipError_t launch_tile_f32(const KArgs<float, XPair> &a, int grid, hipStream_t st) {
  return tile_dispatch(a, grid, st);
}

}  // namespace spx
