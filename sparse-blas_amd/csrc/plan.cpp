// plan.cpp -- device-resident SpMV plans behind the extern "C" ABI.
//
// spx_plan_create uploads A once and prepares everything the kernels need:
//  - rows with very many nonzeros ("long rows") are split off and cut into
//    (column block, row) slices for the segmented engine, which stages each
//    block's x in LDS; k_long_reduce folds the per-block sums;
//  - kernel 4 hands the regular rows to the segmented engine too (one launch);
//  - the remaining "regular" CSR is partitioned on the device (nnz-balanced
//    partitions -> HIP streams; each partition -> workgroup units);
//  - kernel-specific conversion: merge-path unit coordinates, or CSR5-style
//    tiles;
//  - a dry execute records which carry slots belong to which split row (a
//    property of the sparsity structure), giving the fix-up its run table.
// spx_plan_execute then only launches kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/spx_capi.h"
#include "spx_build.h"
#include "spx_internal.h"

#include <thread>

using namespace spx;

namespace {

thread_local std::string g_last_error;

// Test seam: when >= 0, that many more device allocations succeed and the
// next one reports hipErrorOutOfMemory (then the seam disarms).
std::atomic<long long> g_fail_alloc{-1};

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(e_ == hipErrorOutOfMemory ? SPX_EMEM : SPX_EDEV,                     \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                  \
  } while (0)

int default_lanes(double mean) {
  // lanes per row for the CSR-vector kernel: each lane takes two nonzeros
  // per step, so L lanes cover a row of up to 2L - 1 nonzeros (at either
  // alignment) in one step
  if (mean <= 3) return 2;
  if (mean <= 7) return 4;
  if (mean <= 15) return 8;
  if (mean <= 31) return 16;
  if (mean <= 63) return 32;
  return 64;
}

int tile_sigma_auto(int m, int64_t nnz) {
  // SB/include/anonymouslib_cuda.h:297-313
  int64_t q = m > 0 ? nnz / m : 0;
  if (q <= 4) return 4;
  if (q <= 32) return (int)q;
  if (q <= 256) return 32;
  return 6;
}

}  // namespace

struct spx_plan {
  int m = 0, n = 0;
  int64_t nnz = 0;      // whole matrix
  int64_t nnz_reg = 0;  // regular part
  int kernel = 2, mode = SPX_MODE_NNZ, queues = 1;
  int sigma = 0, lanes = 0, items = 0, precision = SPX_PREC_F64;

  int32_t *d_rowptr = nullptr;   // regular CSR row pointer (long rows empty)
  int32_t *d_col = nullptr;      // CSR columns, or tile layout for kernel 3
  double *d_val = nullptr;       // FP64 values (CSR or tile layout)
  float *d_val32 = nullptr;      // FP32 values (precision form)
  VUnit *d_vunits = nullptr;     // kernel 1: unit table
  int32_t *d_vxr = nullptr;      // kernel 1: per launch, kXcd + 1 unit bounds balanced by nonzeros
  std::vector<int64_t> vnnz;     // kernel 1: nonzeros per unit (plan time only)
  MUnit *d_munits = nullptr;     // kernel 2: unit table
  Part *d_parts = nullptr;
  std::vector<Part> parts;
  Unit *d_units = nullptr;
  int64_t nunits = 0;
  Coord *d_coords = nullptr;
  int64_t ncoords = 0;
  uint32_t *d_tp = nullptr, *d_tf = nullptr;
  int64_t ntiles = 0, ntp = 0;
  int32_t *d_crow = nullptr;
  double *d_cval = nullptr;
  int64_t ncarry = 0;
  TUnit *d_tunits = nullptr;     // kernel 3: unit table
  int32_t *d_empty = nullptr;
  int64_t nempty = 0;
  int *d_queue = nullptr;

  // long rows
  int nlong = 0, nblocks = 0;
  int64_t long_nnz = 0, long_th = 0;
  int32_t *d_lrows = nullptr;
  double *d_partial = nullptr;  // one sum per (column block, long row) slice
  int32_t *d_lsp = nullptr, *d_lseg = nullptr;  // long row r's slices: lseg[lsp[r] .. lsp[r+1])
  hipStream_t lstream = nullptr;

  // segmented engine: the long rows' (column block, row) slices for every
  // kernel, plus the regular rows for kernel 4
  SegUnit *d_sunits = nullptr;
  int32_t *d_sxr = nullptr;  // per-XCD unit ranges, balanced by work: all, staged, unstaged (3 x 9)
  uint8_t *d_sstart = nullptr;
  double *d_partk = nullptr;  // multi-RHS scratch: long-row slice partials, column-major
  int64_t npartk = 0;
  bool fused_spmm = false;
  uint32_t *d_scol12 = nullptr;  // staged units' 12-bit column offsets, 3 words per 8 elements
  int32_t *d_scol32 = nullptr;
  double *d_sval = nullptr;
  float *d_sval32 = nullptr;
  int64_t nsunits = 0, nslong = 0;  // units; the first nslong sweep long rows
  int64_t nsstaged = 0;              // split plans: the first nsstaged units stage x
  bool seg_split = false;            // staged and unstaged units run as two concurrent kernels
  int64_t nstaged_units = 0;         // units that stage their x window (any order)
  int64_t nlseg = 0;                 // long-row slices
  bool sstage = false;
  // kernel 4, column bands: the regular rows cut at band edges; slice (b, r)
  // sums into partial[nlseg + k] (band-major, rows ascending within a band)
  int nbands = 0;
  int64_t nbslots = 0;            // band slices
  uint32_t *d_bmask = nullptr;    // per row: bit b = a slice in band b; bit 31 = long row
  int32_t *d_bbase = nullptr;     // [ceil(m/64) * 32]: slot of the first slice of band b at or after row 64w
  hipEvent_t lfork = nullptr, ldone = nullptr;

  struct Launch {
    int stream;
    int64_t ubeg, ucnt;
    int grid;
  };
  std::vector<Launch> launches;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> done;
  hipEvent_t fork = nullptr;
  int64_t bytes = 0;
  double upload_ms = 0.0, build_ms = 0.0;  // host time: copies of the input, the whole build

  ~spx_plan() {
    for (void *p : {(void *)d_rowptr, (void *)d_col, (void *)d_val, (void *)d_val32, (void *)d_vunits, (void *)d_vxr, (void *)d_munits,
                    (void *)d_parts, (void *)d_units, (void *)d_coords, (void *)d_tp, (void *)d_tf,
                    (void *)d_crow, (void *)d_cval, (void *)d_tunits, (void *)d_empty, (void *)d_queue,
                    (void *)d_lrows, (void *)d_partial, (void *)d_sunits, (void *)d_sxr, (void *)d_sstart, (void *)d_partk, (void *)d_lsp, (void *)d_lseg,
                    (void *)d_scol12, (void *)d_scol32, (void *)d_sval, (void *)d_sval32, (void *)d_bmask,
                    (void *)d_bbase})
      if (p) (void)hipFree(p);
    for (auto s : streams) (void)hipStreamDestroy(s);
    for (auto e : done) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (lstream) (void)hipStreamDestroy(lstream);
    if (lfork) (void)hipEventDestroy(lfork);
    if (ldone) (void)hipEventDestroy(ldone);
  }

  template <typename T>
  int alloc(T **p, int64_t count) {
    if (count <= 0) count = 1;
    hipError_t e = dev_malloc((void **)p, sizeof(T) * (size_t)count);
    if (e != hipSuccess) {
      *p = nullptr;
      return fail(e == hipErrorOutOfMemory ? SPX_EMEM : SPX_EDEV,
                  std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    bytes += (int64_t)sizeof(T) * count;
    return SPX_OK;
  }

  template <typename T>
  int upload(T **p, const T *src, int64_t count) {
    int rc = alloc(p, count);
    if (rc) return rc;
    if (count > 0) HIP_TRY(stage_h2d(*p, src, sizeof(T) * (size_t)count));
    return SPX_OK;
  }
};

static int execute_impl(spx_plan *P, double alpha, const double *x, double beta, double *y, hipStream_t st,
                        hipEvent_t t0, hipEvent_t t1, bool dry);

extern "C" {

void spx_plan_desc_init(spx_plan_desc *d) {
  std::memset(d, 0, sizeof(*d));
  d->kernel = SPX_KERNEL_MERGE;
  d->mode = SPX_MODE_NNZ;
  d->parts = 1;
  d->queues = 1;
}

const char *spx_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

void spx::set_last_error(const std::string &msg) { g_last_error = msg; }

// Every device allocation of the library goes through here.
hipError_t spx::dev_malloc(void **p, size_t bytes) {
  long long left = g_fail_alloc.load();
  while (left >= 0) {
    if (g_fail_alloc.compare_exchange_weak(left, left - 1)) {
      if (left == 0) {
        *p = nullptr;
        return hipErrorOutOfMemory;
      }
      break;
    }
  }
  return hipMalloc(p, bytes);
}

extern "C" {
const char *spx_version(void) { return "spx 0.2 gfx950"; }

void spx_debug_fail_alloc(long long after) { g_fail_alloc = after < 0 ? -1 : after; }

}  // extern "C"

// Long-row threshold: rows at least this long are swept by column blocks.
// A row is worth it when it averages at least a few nonzeros per column
// block; below 4096 nonzeros the per-block overhead dominates.
static int64_t long_threshold(const spx_plan_desc *d) {
  if (d->long_row < 0) return -1;
  if (d->long_row > 0) return d->long_row;
  const int64_t blocks = ((int64_t)d->n + kColBlock - 1) / kColBlock;
  return std::max<int64_t>(4096, 4 * blocks);
}

// Host side of the segmented engine: segments in output order, grouped into
// units, nonzeros copied 8-aligned per unit, one start bit per element.
struct SegHost {
  std::vector<SegUnit> units;
  std::vector<int32_t> out, col;  // per segment: output
  std::vector<double> val;
  std::vector<uint8_t> start;
  int64_t pos = 0;  // next free element
  bool stage = false;
};

// Append segments [0, nseg) as units.  seg(i, b, e, o): segment i is source
// elements [b, e) of scol/sval (never empty) with output o; brk(i): a unit
// must not run on into segment i (the caller breaks where the outputs stop
// being consecutive).  A unit takes whole segments while it
// holds at most `target` elements (a longer segment sits alone).  With
// `stage`, a unit whose column window fits kSegStage
// entries stages x in LDS: the builder first grows the unit under that
// window limit and keeps it staged if it reaches half the target (or the
// last segment); otherwise the unit is built unstaged to the full target.
template <typename F, typename G>
static void seg_append(SegHost &H, int64_t nseg, F seg, G brk, const int32_t *scol, const double *sval,
                       int64_t target, bool stage, int flags) {
  auto window = [&](int64_t b, int64_t e, int32_t &lo, int32_t &hi) {
    for (int64_t k = b; k < e; k++) {
      lo = std::min(lo, scol[k]);
      hi = std::max(hi, scol[k]);
    }
  };
  int64_t i = 0;
  while (i < nseg) {
    // extent of a staged unit starting at segment i
    int64_t j = i, el = 0;
    bool staged = false;
    if (stage) {
      int32_t lo = INT32_MAX, hi = -1;
      while (j < nseg) {
        int64_t b, e;
        int32_t o;
        seg(j, b, e, o);
        if (j > i && (el + (e - b) > target || brk(j))) break;
        int32_t l2 = lo, h2 = hi;
        window(b, e, l2, h2);
        if ((int64_t)h2 - l2 + 1 > kSegStage) break;
        lo = l2;
        hi = h2;
        el += e - b;
        j++;
      }
      staged = j > i && (2 * el >= target || j == nseg);
    }
    if (!staged) {  // unstaged: whole segments up to the target
      j = i;
      el = 0;
      while (j < nseg) {
        int64_t b, e;
        int32_t o;
        seg(j, b, e, o);
        if (j > i && (el + (e - b) > target || brk(j))) break;
        el += e - b;
        j++;
      }
    }
    const int64_t E0 = (H.pos + 7) & ~(int64_t)7;
    H.col.resize(E0, 0);
    H.val.resize(E0, 0.0);
    H.start.resize((size_t)((E0 + el + 7) >> 3) + 1, 0);
    SegUnit U = {(int32_t)E0, (int32_t)(E0 + el), (int32_t)H.out.size(), (int32_t)(j - i), 0, 0, 0, flags};
    int32_t lo = INT32_MAX, hi = -1;
    int64_t at = E0;
    for (int64_t q = i; q < j; q++) {
      int64_t b, e;
      int32_t o;
      seg(q, b, e, o);
      window(b, e, lo, hi);
      H.start[(size_t)(at >> 3)] |= (uint8_t)(1u << (at & 7));
      H.out.push_back(o);
      H.col.insert(H.col.end(), scol + b, scol + e);
      H.val.insert(H.val.end(), sval + b, sval + e);
      at += e - b;
    }
    U.o0 = H.out[U.s0];
    U.cmin = hi >= 0 ? lo : 0;
    U.cwin = hi >= 0 ? hi - lo + 1 : 0;
    if (staged) {
      U.flags |= SEG_STAGE;
      H.stage = true;
    }
    H.units.push_back(U);
    H.pos = E0 + el;
    i = j;
  }
}

// The long rows' segments: slice (b, r) = the nonzeros of long row r (columns
// sorted) that fall in column block b, ordered by block, then row, empty
// slices left out; slice k sums into partial[k], and lsp/lseg list each
// row's slices for the reduction.  A unit holds slices of one column block
// only, so it stages one kColBlock-wide window of x in LDS.
static void seg_long(SegHost &H, const std::vector<int32_t> &lrp, const std::vector<int32_t> &lcol,
                     const std::vector<double> &lval, int nlong, int nblocks, int64_t target,
                     std::vector<int32_t> &lsp, std::vector<int32_t> &lseg) {
  static_assert(kColBlock <= kSegStage, "a column block must fit the LDS stage");
  std::vector<int64_t> cut((size_t)nlong * (nblocks + 1));  // cut[r*(nblocks+1)+b]: first element of r in block b
  for (int r = 0; r < nlong; r++) {
    int64_t j = lrp[r];
    for (int b = 0; b <= nblocks; b++) {
      const int64_t lo = (int64_t)b * kColBlock;
      while (j < lrp[r + 1] && lcol[j] < lo) j++;
      cut[(size_t)r * (nblocks + 1) + b] = b == nblocks ? lrp[r + 1] : j;
    }
  }
  std::vector<int64_t> sb, se;
  std::vector<int32_t> srow, sblk;  // long row, column block
  for (int b = 0; b < nblocks; b++)
    for (int r = 0; r < nlong; r++) {
      const int64_t s0 = cut[(size_t)r * (nblocks + 1) + b], s1 = cut[(size_t)r * (nblocks + 1) + b + 1];
      if (s1 <= s0) continue;
      sb.push_back(s0);
      se.push_back(s1);
      srow.push_back(r);
      sblk.push_back(b);
    }
  lsp.assign(nlong + 1, 0);
  for (int32_t r : srow) lsp[r + 1]++;
  for (int r = 0; r < nlong; r++) lsp[r + 1] += lsp[r];
  lseg.assign(srow.size(), 0);
  std::vector<int32_t> fill(lsp.begin(), lsp.end() - 1);
  for (size_t k = 0; k < srow.size(); k++) lseg[fill[srow[k]]++] = (int32_t)k;  // column order per row
  // a block with more than `target` elements is cut into that many nearly
  // equal units (10% slack for whole slices), each staging the block's x
  const int64_t ns = (int64_t)sb.size();
  for (int64_t i = 0; i < ns;) {
    int64_t j = i, tot = 0;
    while (j < ns && sblk[j] == sblk[i]) tot += se[j] - sb[j], j++;
    const int64_t parts = (tot + target - 1) / target;
    const int64_t tb = parts > 1 ? (tot * 11 + parts * 10 - 1) / (parts * 10) : target;
    seg_append(
        H, j - i,
        [&](int64_t q, int64_t &b, int64_t &e, int32_t &o) {
          b = sb[i + q];
          e = se[i + q];
          o = (int32_t)(i + q);
        },
        [](int64_t) { return false; }, lcol.data(), lval.data(), tb, true, SEG_PARTIAL);
    i = j;
  }
}

// Upload the segmented engine's arrays (padded so a group of 8 past the last
// nonzero, and the start byte after it, are readable).
static int seg_upload(spx_plan *P, SegHost &H) {
  const int64_t total = ((H.pos + 7) & ~(int64_t)7) + 8;
  H.col.resize(total, 0);
  H.val.resize(total, 0.0);
  H.start.resize((size_t)(total >> 3), 0);
  // the kernel's preconditions, checked once on the host
  const int64_t nslots = P->nlseg + P->nbslots;
  for (const SegUnit &U : H.units) {
    bool ok = U.e0 % 8 == 0 && U.e0 < U.e1 && U.e1 <= H.pos && U.ns >= 1 && U.s0 >= 0 && U.o0 >= 0 &&
              (int64_t)U.o0 + U.ns <= ((U.flags & SEG_PARTIAL) ? nslots : (int64_t)P->m) &&
              (int64_t)U.s0 + U.ns <= (int64_t)H.out.size() && U.cmin >= 0 && (int64_t)U.cmin + U.cwin <= P->n &&
              (!(U.flags & SEG_STAGE) || U.cwin <= kSegStage) && (H.start[U.e0 >> 3] & 1);
    int64_t starts = 0;
    for (int32_t k = U.e0; ok && k < U.e1; k++) {
      starts += (H.start[k >> 3] >> (k & 7)) & 1;
      ok = H.col[k] >= U.cmin && H.col[k] < U.cmin + U.cwin;
    }
    ok = ok && starts == U.ns;
    for (int s = 0; ok && s < U.ns; s++) ok = H.out[U.s0 + s] == U.o0 + s;  // consecutive outputs
    if (!ok) return fail(SPX_EDEV, "segmented engine: inconsistent unit table");
  }
  // Staged and unstaged units run as two concurrent kernels (staged units
  // ordered first, each class with its own XCD ranges) when the unstaged
  // ones carry a real share of the work and their gathers range over more
  // x than one XCD's L2 holds (4 MiB): those miss L2 anyway and gain from
  // the extra workgroups the kernel without the LDS window fits per CU.
  // Gathers that stay within L2 do better with fewer working sets competing
  // for it, and a few unstaged units ride along in the staged kernel, in
  // their place in the unit order.
  auto work = [](const SegUnit &U) {  // nonzeros, staged x entries, a fixed cost per unit
    return (double)(U.e1 - U.e0) + ((U.flags & SEG_STAGE) ? 0.5 * U.cwin : 0.0) + 256.0;
  };
  {
    double wall = 0.0, wu = 0.0, span = 0.0, el = 0.0;  // el: elements of unstaged units
    int64_t nst = 0;
    for (const SegUnit &U : H.units) {
      wall += work(U);
      if (U.flags & SEG_STAGE) {
        nst++;
      } else {
        wu += work(U);
        span += (double)(U.e1 - U.e0) * U.cwin;
        el += U.e1 - U.e0;
      }
    }
    const double xbytes = 8.0;  // an x entry (FP64, or the FP32 hi/lo pair)
    // column-band plans always split: their slices run without the x window
    P->seg_split = nst > 0 && (P->nbands > 0 || (wu >= 0.1 * wall && el > 0 && span / el * xbytes > 4.0 * (1 << 20)));
    if (P->seg_split)
      std::stable_partition(H.units.begin(), H.units.end(),
                            [](const SegUnit &U) { return (U.flags & SEG_STAGE) != 0; });
    P->nsunits = (int64_t)H.units.size();
    P->nsstaged = P->seg_split ? nst : 0;
    P->nstaged_units = nst;
    P->sstage = nst > 0;
  }
  // XCD ranges of about equal work over all units, and (split plans) over
  // the staged ones and over the rest
  std::vector<double> wsum(H.units.size() + 1, 0.0);
  for (size_t u = 0; u < H.units.size(); u++) wsum[u + 1] = wsum[u] + work(H.units[u]);
  std::vector<int32_t> xr(3 * (kXcd + 1), 0);
  auto ranges = [&](int64_t lo, int64_t hi, int32_t *r) {
    r[0] = (int32_t)lo;
    for (int x = 1; x < kXcd; x++) {
      const double w = wsum[lo] + (wsum[hi] - wsum[lo]) * x / kXcd;
      r[x] = (int32_t)(std::lower_bound(wsum.begin() + lo, wsum.begin() + hi, w) - wsum.begin());
    }
    r[kXcd] = (int32_t)hi;
    for (int x = 1; x <= kXcd; x++) r[x] = std::max(r[x], r[x - 1]);
  };
  ranges(0, P->nsunits, xr.data());
  ranges(0, P->nsstaged, xr.data() + (kXcd + 1));
  ranges(P->nsstaged, P->nsunits, xr.data() + 2 * (kXcd + 1));
  // the device order of the columns and values: values lane-fastest per wave
  // block; columns of staged units as 12-bit window offsets in element
  // order (a lane's 8 are three consecutive words), those of other units in
  // full, lane-fastest, packed unit after unit into col32 from U.cmin on
  std::vector<uint32_t> c12((size_t)(total >> 3) * 3, 0);  // total is a multiple of 8
  std::vector<int32_t> c32;
  {
    std::vector<double> pv((size_t)total, 0.0);
    for (SegUnit &U : H.units) {
      const int64_t e8 = ((int64_t)U.e1 + 7) & ~(int64_t)7;
      for (int64_t e = U.e0; e < e8; e++) pv[(size_t)seg_pos8(e, U.e0, U.e1)] = H.val[(size_t)e];
      if (U.flags & SEG_STAGE) {
        for (int64_t e = U.e0; e < U.e1; e++) {
          const uint32_t v = (uint32_t)(H.col[(size_t)e] - U.cmin);  // < kSegStage <= 4096
          const int64_t bit = (e & 7) * 12;
          const size_t w = (size_t)(e >> 3) * 3 + (size_t)(bit >> 5);
          const int sh = (int)(bit & 31);
          c12[w] |= v << sh;
          if (sh > 20) c12[w + 1] |= v >> (32 - sh);
        }
      } else {
        const size_t base = c32.size();
        if (base + (size_t)(e8 - U.e0) >= ((size_t)1 << 31)) return fail(SPX_ERANGE, "too many unstaged nonzeros");
        c32.resize(base + (size_t)(e8 - U.e0), 0);
        for (int64_t e = U.e0; e < e8; e++)
          c32[base + (size_t)(seg_pos4(e, U.e0, U.e1) - U.e0)] = e < U.e1 ? H.col[(size_t)e] : 0;
        U.cmin = (int32_t)base;
        U.cwin = 0;
      }
    }
    if (P->precision == SPX_PREC_F32) {  // FP32 values follow the 4-byte order
      std::vector<double> p4((size_t)total, 0.0);
      for (const SegUnit &U : H.units) {
        const int64_t e8 = ((int64_t)U.e1 + 7) & ~(int64_t)7;
        for (int64_t e = U.e0; e < e8; e++) p4[(size_t)seg_pos4(e, U.e0, U.e1)] = H.val[(size_t)e];
      }
      pv.swap(p4);
    }
    H.val.swap(pv);
  }
  std::vector<int32_t>().swap(H.col);
  int rc;
  if ((rc = P->upload(&P->d_sunits, H.units.data(), P->nsunits))) return rc;
  if ((rc = P->upload(&P->d_sxr, xr.data(), 3 * (kXcd + 1)))) return rc;
  if ((rc = P->upload(&P->d_sstart, H.start.data(), (int64_t)H.start.size()))) return rc;
  if ((rc = P->upload(&P->d_scol12, c12.data(), (int64_t)c12.size()))) return rc;
  if (!c32.empty() && (rc = P->upload(&P->d_scol32, c32.data(), (int64_t)c32.size()))) return rc;
  if (P->precision == SPX_PREC_F32) {
    std::vector<float> v32(H.val.begin(), H.val.end());
    if ((rc = P->upload(&P->d_sval32, v32.data(), total))) return rc;
  } else {
    if ((rc = P->upload(&P->d_sval, H.val.data(), total))) return rc;
  }
  return SPX_OK;
}

// Device memory for the duration of a plan build (freed on every path).
template <typename T>
struct DevTmp {
  T *p = nullptr;
  DevTmp() = default;
  DevTmp(const DevTmp &) = delete;
  DevTmp &operator=(const DevTmp &) = delete;
  ~DevTmp() {
    if (p) (void)hipFree(p);
  }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  int alloc(int64_t count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    hipError_t e = dev_malloc((void **)&p, sizeof(T) * (size_t)(count > 0 ? count : 1));
    if (e != hipSuccess) {
      p = nullptr;
      return fail(e == hipErrorOutOfMemory ? SPX_EMEM : SPX_EDEV, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return SPX_OK;
  }
};

// The input's columns and values copied to the device (host input of
// kernels 1-3), padded by two zero entries.
struct InputOnDevice {
  DevTmp<int32_t> col;
  DevTmp<double> val;
};

// Kernels 1-3: the long rows' ranges in the input and their own row pointer.
struct LongSplit {
  std::vector<LongRange> lr;  // input range and destination of each long row, row order
  std::vector<int32_t> lrp;   // row pointer of the long rows' arrays
  int64_t ln = 0;             // long-row nonzeros
};

// Fills S from rp and the long rows (ascending), then turns rp in place
// into the row pointer of the regular part, where the long rows are empty.
static void split_long_rows(std::vector<int32_t> &rp, const std::vector<int32_t> &long_rows, LongSplit &S) {
  const int m = (int)rp.size() - 1, nl = (int)long_rows.size();
  S.lr.resize((size_t)nl);
  S.lrp.assign((size_t)nl + 1, 0);
  S.ln = 0;
  for (int j = 0; j < nl; j++) {
    const int r = long_rows[(size_t)j];
    S.lr[(size_t)j] = LongRange{rp[r], rp[r + 1], S.ln};
    S.ln += rp[r + 1] - rp[r];
    S.lrp[(size_t)j + 1] = (int32_t)S.ln;
  }
  if (nl == 0) return;
  size_t j = 0;
  int32_t sub = 0, prev = rp[0];  // prev: the input's rp[i]
  for (int i = 0; i < m; i++) {
    const int32_t next = rp[i + 1];
    if (j < long_rows.size() && long_rows[j] == i) {
      sub += next - prev;
      j++;
    }
    prev = next;
    rp[i + 1] = next - sub;
  }
}

// Kernel 1's lanes per row: the caller's, or from the mean row length.
static int vec_lanes(const spx_plan_desc *d, int64_t nnzr) {
  return d->lanes > 0 ? d->lanes : default_lanes(d->m > 0 ? (double)nnzr / d->m : 0);
}

// Kernel 1's units over rows [lo, hi) of a partition holding nonzeros
// [s0, e1) of a regular part with nnzr nonzeros: the first row of each is
// appended to out.  A unit takes rows in order while it holds at most rcap
// rows (8 passes of the kBlock / lanes row groups, at most kVecUnitRows) and
// cap nonzeros of the partition (kBlock * kVecItems, less for a matrix too
// small to fill about 2,048 such units); a longer row is a unit alone.  Rows
// [r0, r) hold min(rp[r], e1) - max(rp[r0], s0) nonzeros, which grows with
// r, so each unit's end is found by binary search.
static void vec_units(const std::vector<int32_t> &rp, int64_t lo, int64_t hi, int64_t s0, int64_t e1, int lanes,
                      int64_t nnzr, std::vector<int32_t> &out) {
  const int64_t rcap = std::min<int64_t>(kVecUnitRows, 8 * std::max(1, kBlock / lanes));
  const int64_t cap = std::min<int64_t>((int64_t)kBlock * kVecItems, std::max<int64_t>(256, (nnzr + 2047) / 2048));
  for (int64_t r0 = lo; r0 < hi;) {
    out.push_back((int32_t)r0);
    const int64_t rmax = std::min<int64_t>(hi, r0 + rcap);
    const int64_t limit = cap + std::max<int64_t>(rp[r0], s0);
    int64_t r = rmax;
    if (e1 > limit)
      r = (int64_t)(std::upper_bound(rp.begin() + r0 + 1, rp.begin() + rmax + 1,
                                     (int32_t)std::min<int64_t>(limit, INT32_MAX)) -
                    rp.begin()) - 1;
    r0 = r > r0 ? r : r0 + 1;
  }
}

static int build_seg(spx_plan *P, const spx_plan_desc *d, std::vector<int32_t> &rp,
                     const std::vector<int32_t> &long_rows);
static int build_csr(spx_plan *P, const spx_plan_desc *d, std::vector<int32_t> &rp,
                     const std::vector<int32_t> &long_rows, const LongSplit &ls,
                     const std::vector<int32_t> &vwhole, InputOnDevice &in);

static int plan_build(spx_plan *P, const spx_plan_desc *d) {
  const int m = d->m, n = d->n;
  const int64_t nnz = d->nnz;
  if (m < 0 || n < 0 || nnz < 0) return fail(SPX_EARG, "negative size");
  if (nnz >= ((int64_t)1 << 31) - 1) return fail(SPX_ERANGE, "nnz must be below 2^31 per plan");
  if (d->kernel < 1 || d->kernel > 4) return fail(SPX_EARG, "kernel must be 1, 2, 3 or 4");
  if (d->mode < 0 || d->mode > 2) return fail(SPX_EARG, "bad mode");
  if (d->precision != SPX_PREC_F64 && d->precision != SPX_PREC_F32) return fail(SPX_EARG, "bad precision");
  if (m > 0 && !d->rowptr64 && !d->rowptr32) return fail(SPX_EARG, "missing row pointer");
  if (nnz > 0 && (!d->col || !d->val)) return fail(SPX_EARG, "missing col/val");
  if (d->mode == SPX_MODE_QUEUE && d->nb <= 0) return fail(SPX_EARG, "nb must be positive");
  P->m = m;
  P->n = n;
  P->nnz = nnz;
  P->kernel = d->kernel;
  P->mode = d->mode;
  P->precision = d->precision;
  P->queues = d->queues > 0 ? d->queues : 1;
  P->fused_spmm = (d->flags & SPX_FLAG_FUSED_SPMM) != 0;
  if (m == 0) return SPX_OK;
  const bool on_dev = (d->flags & SPX_FLAG_DEVICE_PTRS) != 0;

  // ---- kernels 1-3 from host arrays: the columns and values go to the
  // device on a second host thread while this one works on the row pointer
  // (everything the upload thread touches is declared before the guard that
  // joins it, so an early return joins the thread before any of it dies)
  InputOnDevice in;
  std::thread upload;
  int up_rc = SPX_OK;
  std::string up_msg;
  struct Join {
    std::thread &t;
    ~Join() {
      if (t.joinable()) t.join();
    }
  } join{upload};
  if (d->kernel != SPX_KERNEL_SEG && !on_dev && nnz > 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    const auto t_up = std::chrono::steady_clock::now();
    upload = std::thread([&, dev, t_up] {
      // a bad row pointer found meanwhile takes precedence over whatever
      // happens here (this thread's error text is handed over on return)
      int cur = -1, rc = SPX_OK;
      if (hipGetDevice(&cur) == hipSuccess && cur != dev && hipSetDevice(dev) != hipSuccess)
        rc = fail(SPX_EDEV, "upload: hipSetDevice failed");
      if (!rc) rc = in.col.alloc(nnz + 2);
      if (!rc) rc = in.val.alloc(nnz + 2);
      hipError_t e = hipSuccess;
      if (!rc) e = hipMemset(in.col.p + nnz, 0, 2 * sizeof(int32_t));
      if (!rc && e == hipSuccess) e = hipMemset(in.val.p + nnz, 0, 2 * sizeof(double));
      if (!rc && e == hipSuccess) e = stage_h2d(in.col.p, d->col, sizeof(int32_t) * (size_t)nnz);
      if (!rc && e == hipSuccess) e = stage_h2d(in.val.p, d->val, sizeof(double) * (size_t)nnz);
      if (!rc && e != hipSuccess)
        rc = fail(e == hipErrorOutOfMemory ? SPX_EMEM : SPX_EDEV, std::string("upload: ") + hipGetErrorString(e));
      if (rc) up_msg = spx_last_error();
      up_rc = rc;
      P->upload_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_up).count();
    });
  }

  // ---- host row pointer (validated: monotone, from 0 to nnz) ----
  std::vector<int64_t> rp64;
  std::vector<int32_t> rp32;
  const int64_t *src64 = d->rowptr64;
  const int32_t *src32 = d->rowptr32;
  if (on_dev) {
    if (src64) {
      rp64.resize(m + 1);
      HIP_TRY(hipMemcpy(rp64.data(), src64, sizeof(int64_t) * (m + 1), hipMemcpyDeviceToHost));
      src64 = rp64.data();
    } else {
      rp32.resize(m + 1);
      HIP_TRY(hipMemcpy(rp32.data(), src32, sizeof(int32_t) * (m + 1), hipMemcpyDeviceToHost));
      src32 = rp32.data();
    }
  }
  std::vector<int32_t> rp(m + 1);
  for (int i = 0; i <= m; i++) {
    int64_t v = src64 ? src64[i] : (int64_t)src32[i];
    if (v < 0 || v > nnz || (i > 0 && v < rp[i - 1])) return fail(SPX_EARG, "row pointer not monotone within [0, nnz]");
    rp[i] = (int32_t)v;
  }
  if (rp[0] != 0 || rp[m] != nnz) return fail(SPX_EARG, "row pointer must start at 0 and end at nnz");

  // ---- split off long rows ----
  P->long_th = long_threshold(d);
  std::vector<int32_t> long_rows;
  if (P->long_th > 0) {
    // at most kMaxLong long rows (the longest ones): raise the threshold if needed
    std::vector<int32_t> lens;
    for (int i = 0; i < m; i++)
      if (rp[i + 1] - rp[i] >= P->long_th) lens.push_back(rp[i + 1] - rp[i]);
    if ((int)lens.size() > kMaxLong) {
      std::nth_element(lens.begin(), lens.begin() + kMaxLong, lens.end(), std::greater<int32_t>());
      P->long_th = (int64_t)lens[kMaxLong] + 1;
    }
    for (int i = 0; i < m; i++)
      if (rp[i + 1] - rp[i] >= P->long_th) long_rows.push_back(i);
    if ((int)long_rows.size() > kMaxLong) long_rows.resize(kMaxLong);  // ties at the cut
  }
  P->nlong = (int)long_rows.size();
  // kernels 1-3: the regular row pointer, while the upload may still run;
  // kernel 1 also cuts the units of a plan with one partition (all rows),
  // while the row pointer is in the host's caches
  LongSplit ls;
  std::vector<int32_t> vwhole;
  if (d->kernel != SPX_KERNEL_SEG) split_long_rows(rp, long_rows, ls);
  if (d->kernel == SPX_KERNEL_VECTOR) {
    const int lanes = vec_lanes(d, nnz - ls.ln);
    if (lanes >= 1 && lanes <= 64 && !(lanes & (lanes - 1))) vec_units(rp, 0, m, 0, nnz - ls.ln, lanes, nnz - ls.ln, vwhole);
  }

  if (upload.joinable()) upload.join();
  if (up_rc != SPX_OK) return fail(up_rc, up_msg);
  if (d->kernel == SPX_KERNEL_SEG) return build_seg(P, d, rp, long_rows);
  return build_csr(P, d, rp, long_rows, ls, vwhole, in);
}

// Column bands for kernel 4's regular rows.  A unit of rows whose columns
// spread over more of x than one XCD's L2 holds gathers every entry from the
// MALL or HBM, a whole cache line per 8-byte entry.  When most regular
// nonzeros sit in rows wider than one band, the rows are cut at band edges
// instead: band b's slices (rows ascending) form their own units, so the
// units an XCD runs at a time gather from one band of x that stays in its
// L2, and each slice's sum goes to a partial slot; k_band_reduce then adds a
// row's slices in band order.  Bands are kBandCols columns (2 MiB of x) wide,
// at most 31 per plan (wider when x is larger).
constexpr int64_t kBandCols = 1 << 18;
constexpr int kMaxBands = 31;

// The number of bands for these regular rows, or 0 for the row-unit layout.
static int band_count(int m, int n, const std::vector<int32_t> &rp, const int32_t *col) {
  if ((int64_t)n < 2 * kBandCols) return 0;
  int64_t wide = 0;
  for (int r = 0; r < m; r++) {
    const int64_t b = rp[r], e = rp[r + 1];
    if (e - b < 2) continue;
    int32_t lo = col[b], hi = col[b];
    for (int64_t k = b + 1; k < e; k++) {
      lo = std::min(lo, col[k]);
      hi = std::max(hi, col[k]);
    }
    if ((int64_t)hi - lo >= kBandCols) wide += e - b;
  }
  if (2 * wide < (int64_t)rp[m]) return 0;
  return (int)std::min<int64_t>(kMaxBands, ((int64_t)n + kBandCols - 1) / kBandCols);
}

// The regular rows as band slices in units (SEG_PARTIAL), plus the per-row
// band masks and per-64-row slot bases for k_band_reduce.
static int seg_bands(spx_plan *P, SegHost &H, int nb, const std::vector<int32_t> &rp, const int32_t *col,
                     const double *val, const std::vector<int32_t> &long_rows, int64_t target) {
  const int m = P->m;
  const int64_t width = ((int64_t)P->n + nb - 1) / nb;
  const int64_t nnz = rp[m];
  std::vector<uint32_t> mask((size_t)m, 0);
  for (int r : long_rows) mask[r] = 1u << 31;
  std::vector<int64_t> elems(nb + 1, 0), slices(nb + 1, 0);
  for (int r = 0; r < m; r++) {
    uint32_t mk = 0;
    for (int64_t k = rp[r]; k < rp[r + 1]; k++) {
      const int b = (int)(col[k] / width);
      mk |= 1u << b;
      elems[b + 1]++;
    }
    for (uint32_t t = mk; t; t &= t - 1) slices[__builtin_ctz(t) + 1]++;
    mask[r] |= mk;
  }
  for (int b = 0; b < nb; b++) {
    elems[b + 1] += elems[b];
    slices[b + 1] += slices[b];
  }
  const int64_t S = slices[nb];
  if (P->nlseg + S >= INT32_MAX) return fail(SPX_ERANGE, "too many band slices");
  // band-major copies of the columns and values, each slice's range, and
  // the slot of each band's first slice at every 64th row
  std::vector<int32_t> bcol((size_t)nnz);
  std::vector<double> bval((size_t)nnz);
  std::vector<int64_t> sb((size_t)S);
  std::vector<int32_t> sl((size_t)S);
  std::vector<int64_t> epos(elems.begin(), elems.end() - 1), spos(slices.begin(), slices.end() - 1);
  const int64_t nw = ((int64_t)m + kWave - 1) / kWave;
  std::vector<int32_t> base((size_t)(nw * (kMaxBands + 1)), 0);  // 32 per wave, unused bands 0
  for (int r = 0; r < m; r++) {
    if (r % kWave == 0)
      for (int b = 0; b < nb; b++) base[(size_t)(r / kWave) * (kMaxBands + 1) + b] = (int32_t)(P->nlseg + spos[b]);
    const uint32_t mk = mask[r] & ~(1u << 31);
    for (uint32_t t = mk; t; t &= t - 1) {
      const int b = __builtin_ctz(t);
      sb[(size_t)spos[b]] = epos[b];
    }
    for (int64_t k = rp[r]; k < rp[r + 1]; k++) {
      const int b = (int)(col[k] / width);
      bcol[(size_t)epos[b]] = col[k];
      bval[(size_t)epos[b]] = val[k];
      epos[b]++;
    }
    for (uint32_t t = mk; t; t &= t - 1) {
      const int b = __builtin_ctz(t);
      sl[(size_t)spos[b]] = (int32_t)(epos[b] - sb[(size_t)spos[b]]);
      spos[b]++;
    }
  }
  for (int b = 0; b < nb; b++) {
    const int64_t s0 = slices[b];
    seg_append(
        H, slices[b + 1] - s0,
        [&](int64_t q, int64_t &e0, int64_t &e1, int32_t &o) {
          e0 = sb[(size_t)(s0 + q)];
          e1 = e0 + sl[(size_t)(s0 + q)];
          o = (int32_t)(P->nlseg + s0 + q);
        },
        [](int64_t) { return false; }, bcol.data(), bval.data(), target, true, SEG_PARTIAL);
    if (H.pos + 64 >= INT32_MAX) return fail(SPX_ERANGE, "padded nonzeros must stay below 2^31 per plan");
  }
  P->nbands = nb;
  P->nbslots = S;
  int rc;
  if ((rc = P->upload(&P->d_bmask, mask.data(), m))) return rc;
  if ((rc = P->upload(&P->d_bbase, base.data(), nw * (kMaxBands + 1)))) return rc;
  return SPX_OK;
}

// Kernel 4: the unit table is built on the host (it needs each unit's column
// window), then uploaded in the segmented layout.
static int build_seg(spx_plan *P, const spx_plan_desc *d, std::vector<int32_t> &rp,
                     const std::vector<int32_t> &long_rows) {
  const int m = d->m, n = d->n;
  const int64_t nnz = d->nnz;
  const bool on_dev = (d->flags & SPX_FLAG_DEVICE_PTRS) != 0;
  const int32_t *src_col = d->col;
  const double *src_val = d->val;
  std::vector<int32_t> hcol, rcol, lcol, lrp;
  std::vector<double> hval, rval, lval;
  if (on_dev) {
    hcol.resize(nnz);
    hval.resize(nnz);
    HIP_TRY(hipMemcpy(hcol.data(), d->col, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hval.data(), d->val, sizeof(double) * nnz, hipMemcpyDeviceToHost));
    src_col = hcol.data();
    src_val = hval.data();
  }
  for (int64_t k = 0; k < nnz; k++)
    if (src_col[k] < 0 || src_col[k] >= n) return fail(SPX_EARG, "column index out of range");
  if (P->nlong > 0) {
    std::vector<char> is_long(m, 0);
    for (int r : long_rows) is_long[r] = 1;
    int64_t ln = 0;
    for (int r : long_rows) ln += rp[r + 1] - rp[r];
    P->long_nnz = ln;
    rcol.reserve(nnz - ln);
    rval.reserve(nnz - ln);
    lcol.reserve(ln);
    lval.reserve(ln);
    lrp.push_back(0);
    std::vector<int32_t> rrp(m + 1, 0);
    std::vector<std::pair<int32_t, double>> tmp;
    for (int i = 0; i < m; i++) {
      const int64_t b = rp[i], e = rp[i + 1];
      if (is_long[i]) {
        // columns sorted within the row (a stable sort keeps equal columns in order)
        bool sorted = true;
        for (int64_t j = b + 1; j < e && sorted; j++) sorted = src_col[j - 1] <= src_col[j];
        if (sorted) {
          lcol.insert(lcol.end(), src_col + b, src_col + e);
          lval.insert(lval.end(), src_val + b, src_val + e);
        } else {
          tmp.clear();
          for (int64_t j = b; j < e; j++) tmp.emplace_back(src_col[j], src_val[j]);
          std::stable_sort(tmp.begin(), tmp.end(),
                           [](const std::pair<int32_t, double> &a, const std::pair<int32_t, double> &c) {
                             return a.first < c.first;
                           });
          for (auto &t : tmp) {
            lcol.push_back(t.first);
            lval.push_back(t.second);
          }
        }
        lrp.push_back((int32_t)lcol.size());
      } else {
        rcol.insert(rcol.end(), src_col + b, src_col + e);
        rval.insert(rval.end(), src_val + b, src_val + e);
      }
      rrp[i + 1] = (int32_t)rcol.size();
    }
    rp.swap(rrp);
    src_col = rcol.data();
    src_val = rval.data();
  }
  const int64_t nnzr = rp[m];
  P->nnz_reg = nnzr;

  // ---- long rows: (column block, row) slices for the segmented engine ----
  SegHost H;
  const int64_t seg_target = (int64_t)kSegBlock * 8 * (d->kernel == SPX_KERNEL_SEG && d->items > 0 ? d->items : 2);
  if (d->kernel == SPX_KERNEL_SEG && (d->items < 0 || d->items > 32))
    return fail(SPX_EARG, "items (chunks per unit) must be in 1..32");
  if (P->nlong > 0) {
    P->nblocks = (int)(((int64_t)n + kColBlock - 1) / kColBlock);
    if ((int64_t)P->nblocks * P->nlong >= INT32_MAX) return fail(SPX_ERANGE, "too many long-row blocks");
    if (P->long_nnz + 8 * (int64_t)P->nblocks * P->nlong + 64 >= INT32_MAX)
      return fail(SPX_ERANGE, "padded long-row nonzeros must stay below 2^31 per plan");
    std::vector<int32_t> lsp, lseg;
    seg_long(H, lrp, lcol, lval, P->nlong, P->nblocks, std::max<int64_t>(seg_target, kLongUnit), lsp, lseg);
    P->nslong = (int64_t)H.units.size();
    P->nlseg = (int64_t)lseg.size();
    int rc;
    if ((rc = P->upload(&P->d_lrows, long_rows.data(), P->nlong))) return rc;
    if ((rc = P->upload(&P->d_lsp, lsp.data(), P->nlong + 1))) return rc;
    if ((rc = P->upload(&P->d_lseg, lseg.data(), P->nlseg))) return rc;
  }

  // ---- the regular rows join the long slices in one launch ----
  {
    P->items = (int)(seg_target / (kSegBlock * 8));
    // empty rows of A (the long rows are written by the reduction)
    std::vector<char> is_long(m, 0);
    for (int r : long_rows) is_long[r] = 1;
    std::vector<int32_t> rows, empty;
    for (int i = 0; i < m; i++)
      if (rp[i + 1] > rp[i])
        rows.push_back(i);
      else if (!is_long[i])
        empty.push_back(i);
    if (P->long_nnz + nnzr + 8 * ((int64_t)rows.size() + (int64_t)P->nblocks * P->nlong) + 64 >= INT32_MAX)
      return fail(SPX_ERANGE, "padded nonzeros must stay below 2^31 per plan");
    int rc;
    const int bands = band_count(m, n, rp, src_col);
    if (bands > 0) {
      if ((rc = seg_bands(P, H, bands, rp, src_col, src_val, long_rows, seg_target))) return rc;
      empty.clear();  // the band reduction writes every regular row
    } else {
      seg_append(
          H, (int64_t)rows.size(),
          [&](int64_t i, int64_t &b, int64_t &e, int32_t &o) {
            o = rows[i];
            b = rp[o];
            e = rp[o + 1];
          },
          [&](int64_t i) { return rows[i] != rows[i - 1] + 1; }, src_col, src_val, seg_target, true, 0);
    }
    if ((rc = seg_upload(P, H))) return rc;
    if (P->nlseg + P->nbslots > 0 && (rc = P->alloc(&P->d_partial, P->nlseg + P->nbslots))) return rc;
    if (P->seg_split) {
      HIP_TRY(hipStreamCreateWithFlags(&P->lstream, hipStreamNonBlocking));
      HIP_TRY(hipEventCreateWithFlags(&P->lfork, hipEventDisableTiming));
      HIP_TRY(hipEventCreateWithFlags(&P->ldone, hipEventDisableTiming));
    }
    P->nempty = (int64_t)empty.size();
    if ((rc = P->upload(&P->d_empty, empty.data(), P->nempty))) return rc;
    HIP_TRY(hipDeviceSynchronize());
    return SPX_OK;
  }
}

// The long rows' (column block, row) slices in the segmented engine's
// layout, built on the device from the long rows' compact arrays (columns
// sorted within each row).  Slices are ordered by block, then row; a block
// is one unit unless it holds more than `target` nonzeros, then it is cut
// into nearly equal units of whole slices.  Every unit stages its block of
// x.  Slice k sums into partial[k]; lsp/lseg list each row's slices in
// column order for k_long_reduce.
static int build_long_slices(spx_plan *P, int n, int32_t *lcol, double *lval, const std::vector<int32_t> &lrp,
                             int64_t target) {
  const int nlong = P->nlong;
  const int nb = P->nblocks;
  int rc;
  DevTmp<int32_t> d_lrp, d_cut, d_spos, d_ssrc;
  DevTmp<int> d_flag;
  if ((rc = d_lrp.alloc(nlong + 1)) || (rc = d_flag.alloc(nlong))) return rc;
  HIP_TRY(hipMemcpy(d_lrp.p, lrp.data(), sizeof(int32_t) * (nlong + 1), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(d_flag.p, 0, sizeof(int) * nlong));
  HIP_TRY(launch_long_unsorted(lcol, d_lrp.p, nlong, d_flag.p, 0));
  std::vector<int> unsorted(nlong);
  HIP_TRY(hipMemcpy(unsorted.data(), d_flag.p, sizeof(int) * nlong, hipMemcpyDeviceToHost));
  for (int r = 0; r < nlong; r++) {
    if (!unsorted[r]) continue;
    // rare: a long row with unsorted columns is sorted on the host (stable,
    // so equal columns keep their order)
    const int64_t b = lrp[r], len = lrp[r + 1] - lrp[r];
    std::vector<int32_t> c(len);
    std::vector<double> v(len);
    HIP_TRY(hipMemcpy(c.data(), lcol + b, sizeof(int32_t) * len, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(v.data(), lval + b, sizeof(double) * len, hipMemcpyDeviceToHost));
    std::vector<int64_t> ord(len);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return c[x] < c[y]; });
    std::vector<int32_t> c2(len);
    std::vector<double> v2(len);
    for (int64_t k = 0; k < len; k++) {
      c2[k] = c[ord[k]];
      v2[k] = v[ord[k]];
    }
    HIP_TRY(hipMemcpy(lcol + b, c2.data(), sizeof(int32_t) * len, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(lval + b, v2.data(), sizeof(double) * len, hipMemcpyHostToDevice));
  }
  const int64_t per = (int64_t)nb + 1;
  if ((rc = d_cut.alloc(per * nlong))) return rc;
  HIP_TRY(launch_long_cuts(lcol, d_lrp.p, nlong, nb, kColBlock, d_cut.p, 0));
  std::vector<int32_t> cut((size_t)(per * nlong));
  HIP_TRY(hipMemcpy(cut.data(), d_cut.p, sizeof(int32_t) * cut.size(), hipMemcpyDeviceToHost));

  // slices, block-major with rows ascending in a block: counted, then
  // placed, walking each row's cuts in order
  std::vector<int64_t> bstart(nb + 1, 0);  // slices of block b: [bstart[b], bstart[b + 1])
  for (int r = 0; r < nlong; r++) {
    const int32_t *c = cut.data() + (size_t)r * per;
    for (int b = 0; b < nb; b++) bstart[b + 1] += c[b + 1] > c[b] ? 1 : 0;
  }
  for (int b = 0; b < nb; b++) bstart[b + 1] += bstart[b];
  const int64_t ns = bstart[nb];
  std::vector<int32_t> sb((size_t)ns), slen((size_t)ns), srow((size_t)ns);
  {
    std::vector<int64_t> at(bstart.begin(), bstart.end() - 1);
    for (int r = 0; r < nlong; r++) {
      const int32_t *c = cut.data() + (size_t)r * per;
      for (int b = 0; b < nb; b++) {
        if (c[b + 1] <= c[b]) continue;
        const size_t k = (size_t)at[b]++;
        sb[k] = c[b];
        slen[k] = c[b + 1] - c[b];
        srow[k] = r;
      }
    }
  }
  P->nlseg = ns;
  std::vector<int32_t> lsp(nlong + 1, 0), lseg((size_t)ns);
  for (int32_t r : srow) lsp[r + 1]++;
  for (int r = 0; r < nlong; r++) lsp[r + 1] += lsp[r];
  {
    std::vector<int32_t> fill(lsp.begin(), lsp.end() - 1);
    for (int64_t k = 0; k < ns; k++) lseg[fill[srow[k]]++] = (int32_t)k;  // column order per row
  }
  // units
  std::vector<SegUnit> units;
  std::vector<int32_t> spos((size_t)ns);
  int64_t pos = 0;
  for (int b = 0; b < nb; b++) {
    const int64_t i0 = bstart[b], i1 = bstart[b + 1];
    int64_t tot = 0;
    for (int64_t k = i0; k < i1; k++) tot += slen[k];
    if (tot == 0) continue;
    const int64_t parts = (tot + target - 1) / target;
    const int64_t tb = parts > 1 ? (tot * 11 + parts * 10 - 1) / (parts * 10) : target;
    const int32_t cmin = b * kColBlock, cwin = std::min<int32_t>(kColBlock, n - cmin);
    for (int64_t k = i0; k < i1;) {
      int64_t el = 0, q = k;
      while (q < i1 && (q == k || el + slen[q] <= tb)) el += slen[q++];
      const int64_t E0 = (pos + 7) & ~(int64_t)7;
      if (E0 + el + 16 >= INT32_MAX) return fail(SPX_ERANGE, "padded long-row nonzeros must stay below 2^31");
      units.push_back(SegUnit{(int32_t)E0, (int32_t)(E0 + el), (int32_t)k, (int32_t)(q - k), (int32_t)k, cmin, cwin,
                              SEG_STAGE | SEG_PARTIAL});
      int64_t at = E0;
      for (int64_t t = k; t < q; t++) {
        spos[t] = (int32_t)at;
        at += slen[t];
      }
      pos = E0 + el;
      k = q;
    }
  }
  const int64_t nu = (int64_t)units.size();
  const int64_t total = ((pos + 7) & ~(int64_t)7) + 8, groups = total / 8;
  P->nsunits = P->nslong = P->nstaged_units = nu;
  P->nsstaged = 0;
  P->seg_split = false;
  P->sstage = nu > 0;
  // XCD ranges of about equal work (nonzeros, staged x entries, a fixed cost per unit)
  std::vector<double> wsum((size_t)nu + 1, 0.0);
  for (int64_t u = 0; u < nu; u++)
    wsum[u + 1] = wsum[u] + (double)(units[u].e1 - units[u].e0) + 0.5 * units[u].cwin + 256.0;
  std::vector<int32_t> xr(3 * (kXcd + 1), 0);
  xr[0] = 0;
  for (int x = 1; x < kXcd; x++)
    xr[x] = (int32_t)(std::lower_bound(wsum.begin(), wsum.end() - 1, wsum[nu] * x / kXcd) - wsum.begin());
  xr[kXcd] = (int32_t)nu;
  for (int x = 1; x <= kXcd; x++) xr[x] = std::max(xr[x], xr[x - 1]);
  for (int x = 0; x <= kXcd; x++) xr[2 * (kXcd + 1) + x] = xr[x];  // "unstaged" class: the same units
  if ((rc = P->upload(&P->d_sunits, units.data(), nu))) return rc;
  if ((rc = P->upload(&P->d_sxr, xr.data(), 3 * (kXcd + 1)))) return rc;
  if ((rc = d_spos.alloc(ns)) || (rc = d_ssrc.alloc(ns))) return rc;
  if (ns) {
    HIP_TRY(hipMemcpy(d_spos.p, spos.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_ssrc.p, sb.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice));
  }
  if ((rc = P->alloc(&P->d_sstart, groups)) || (rc = P->alloc(&P->d_scol12, 3 * groups))) return rc;
  const bool f32 = P->precision == SPX_PREC_F32;
  if (f32) {
    if ((rc = P->alloc(&P->d_sval32, total))) return rc;
    HIP_TRY(hipMemset(P->d_sval32, 0, sizeof(float) * total));
  } else {
    if ((rc = P->alloc(&P->d_sval, total))) return rc;
    HIP_TRY(hipMemset(P->d_sval, 0, sizeof(double) * total));
  }
  HIP_TRY(launch_seg_fill(P->d_sunits, nu, d_spos.p, d_ssrc.p, lcol, lval, groups, P->d_sstart, P->d_scol12,
                          f32 ? nullptr : P->d_sval, f32 ? P->d_sval32 : nullptr, 0));
  if ((rc = P->upload(&P->d_lsp, lsp.data(), nlong + 1))) return rc;
  if ((rc = P->upload(&P->d_lseg, lseg.data(), ns))) return rc;
  if ((rc = P->alloc(&P->d_partial, ns))) return rc;
  HIP_TRY(hipDeviceSynchronize());
  return SPX_OK;
}

// Kernels 1-3: the input reaches device memory in one copy per array and
// the conversion runs on the device (long rows split off and laid out for
// the segmented engine, columns checked), so a one-shot call costs the
// upload and a few passes at HBM speed; the host touches only the row
// pointer.
static int build_csr(spx_plan *P, const spx_plan_desc *d, std::vector<int32_t> &rp,
                     const std::vector<int32_t> &long_rows, const LongSplit &ls,
                     const std::vector<int32_t> &vwhole, InputOnDevice &in) {
  const int m = d->m, n = d->n;
  const int64_t nnz = d->nnz;
  const bool on_dev = (d->flags & SPX_FLAG_DEVICE_PTRS) != 0;
  // FP32 values in CSR order (kernel 3 converts while it builds its tiles)
  const bool f32_csr = P->precision == SPX_PREC_F32 && d->kernel != SPX_KERNEL_TILE;
  int rc;

  // long rows split off (rp is already the regular part's row pointer)
  const std::vector<LongRange> &lr = ls.lr;
  const std::vector<int32_t> &lrp = ls.lrp;
  const int64_t ln = ls.ln;
  P->long_nnz = ln;
  const int64_t nnzr = nnz - ln;
  P->nnz_reg = nnzr;

  // ---- the regular CSR on the device, padded by two zero entries ----
  // the input's columns and values in device memory: the caller's (device
  // input) or the copies made while the row pointer was read (host input)
  const int32_t *icol = on_dev ? d->col : in.col.p;
  const double *ival = on_dev ? d->val : in.val.p;
  DevTmp<int> d_bad;
  if ((rc = d_bad.alloc(1))) return rc;
  HIP_TRY(hipMemset(d_bad.p, 0, sizeof(int)));
  if (nnz > 0) HIP_TRY(launch_check_cols(icol, nnz, n, d_bad.p, 0));
  DevTmp<int32_t> lcol;
  DevTmp<double> lval;
  if (P->nlong == 0 && in.col.p) {
    // the copies become the plan's arrays
    P->d_col = in.col.p;
    in.col.p = nullptr;
    P->bytes += (int64_t)sizeof(int32_t) * (nnzr + 2);
    if (!f32_csr) {
      P->d_val = in.val.p;
      in.val.p = nullptr;
      P->bytes += (int64_t)sizeof(double) * (nnzr + 2);
    }
  } else {
    if ((rc = P->alloc(&P->d_col, nnzr + 2))) return rc;
    HIP_TRY(hipMemset(P->d_col + nnzr, 0, 2 * sizeof(int32_t)));
  }
  if (!P->d_val) {
    if (f32_csr) {
      if ((rc = P->alloc(&P->d_val32, nnzr + 2))) return rc;
      HIP_TRY(hipMemset(P->d_val32 + nnzr, 0, 2 * sizeof(float)));
    } else {
      if ((rc = P->alloc(&P->d_val, nnzr + 2))) return rc;
      HIP_TRY(hipMemset(P->d_val + nnzr, 0, 2 * sizeof(double)));
    }
  }
  if (P->nlong == 0) {
    if (nnz > 0) {
      if (icol != P->d_col) HIP_TRY(hipMemcpy(P->d_col, icol, sizeof(int32_t) * nnz, hipMemcpyDeviceToDevice));
      if (f32_csr)
        HIP_TRY(launch_to_f32(ival, P->d_val32, nnz, 0));
      else if (ival != P->d_val)
        HIP_TRY(hipMemcpy(P->d_val, ival, sizeof(double) * nnz, hipMemcpyDeviceToDevice));
    }
  } else {
    DevTmp<LongRange> d_lr;
    if ((rc = d_lr.alloc(P->nlong)) || (rc = lcol.alloc(ln)) || (rc = lval.alloc(ln))) return rc;
    HIP_TRY(hipMemcpy(d_lr.p, lr.data(), sizeof(LongRange) * lr.size(), hipMemcpyHostToDevice));
    HIP_TRY(launch_split_long(icol, ival, nnz, d_lr.p, P->nlong, P->d_col, f32_csr ? nullptr : P->d_val,
                              f32_csr ? P->d_val32 : nullptr, lcol.p, lval.p, 0));
  }
  HIP_TRY(hipDeviceSynchronize());
  {
    int bad = 0;
    HIP_TRY(hipMemcpy(&bad, d_bad.p, sizeof(int), hipMemcpyDeviceToHost));
    if (bad) return fail(SPX_EARG, "column index out of range");
  }
  in.col.reset();
  in.val.reset();

  // ---- long rows: (column block, row) slices for the segmented engine ----
  if (P->nlong > 0) {
    P->nblocks = (int)(((int64_t)n + kColBlock - 1) / kColBlock);
    if ((int64_t)P->nblocks * P->nlong >= INT32_MAX) return fail(SPX_ERANGE, "too many long-row blocks");
    if (P->long_nnz + 8 * (int64_t)P->nblocks * P->nlong + 64 >= INT32_MAX)
      return fail(SPX_ERANGE, "padded long-row nonzeros must stay below 2^31 per plan");
    if ((rc = build_long_slices(P, n, lcol.p, lval.p, lrp, std::max<int64_t>((int64_t)kSegBlock * 8 * 2, kLongUnit))))
      return rc;
    if ((rc = P->upload(&P->d_lrows, long_rows.data(), P->nlong))) return rc;
  }
  lcol.reset();
  lval.reset();

  // ---- partitions of the regular part ----
  int kind = PART_NNZ, np = 1;
  if (d->mode == SPX_MODE_ROWS) {
    kind = PART_ROWS;
    np = std::max(1, std::min(d->parts, m));
  } else if (d->mode == SPX_MODE_NNZ) {
    kind = PART_NNZ;
    np = std::max(1, d->parts);
    if (nnzr > 0 && np > nnzr) np = (int)nnzr;
    if (nnzr == 0) np = 1;
  } else {
    kind = PART_TASKS;
    int64_t T = nnzr > 0 ? (nnzr + d->nb - 1) / d->nb : 1;  // SB/src/dspmv_mgpu_v2.cu:218
    if (T > (1 << 24)) return fail(SPX_ERANGE, "too many tasks (nb too small)");
    np = (int)T;
  }

  if ((rc = P->upload(&P->d_rowptr, rp.data(), m + 1))) return rc;
  if ((rc = P->alloc(&P->d_parts, np))) return rc;
  if (nnzr == 0) {
    // every regular row empty: one partition owning all rows
    Part q = {};
    q.s = 0;
    q.e = -1;
    q.row_lo = 0;
    q.row_hi = m;
    P->parts.assign(1, q);
  } else {
    HIP_TRY(launch_partition(P->d_rowptr, m, nnzr, np, kind, P->d_parts, 0));
    P->parts.resize(np);
    HIP_TRY(hipMemcpy(P->parts.data(), P->d_parts, sizeof(Part) * np, hipMemcpyDeviceToHost));
  }

  // ---- per-kernel work decomposition ----
  if (P->kernel == SPX_KERNEL_VECTOR) {
    P->lanes = vec_lanes(d, nnzr);
    if (P->lanes > 64 || (P->lanes & (P->lanes - 1))) return fail(SPX_EARG, "lanes must be a power of two <= 64");
  } else if (P->kernel == SPX_KERNEL_MERGE) {
    P->items = d->items > 0 ? d->items : 8;
    const int ok[] = {2, 4, 6, 8, 12, 16};
    if (std::find(std::begin(ok), std::end(ok), P->items) == std::end(ok))
      return fail(SPX_EARG, "items must be one of 2,4,6,8,12,16");
  } else {
    P->sigma = d->sigma > 0 ? d->sigma : tile_sigma_auto(m, nnzr);
    if (P->sigma < 1 || P->sigma > 32) return fail(SPX_EARG, "sigma must be in 1..32");
  }
  // units never cross partitions; every unit u (global numbering, in
  // nonzero order) owns carry slots 2u and 2u + 1
  const int tw = kBlock * P->items;
  // kernel 3: up to kTileUnitChunks chunks of 4 tiles per unit, fewer when the
  // matrix has too few tiles to give about 1,024 units
  int64_t tiles_per_unit = 4 * kTileUnitChunks;
  if (P->kernel == SPX_KERNEL_TILE) {
    int64_t tiles = 0;
    for (const auto &q : P->parts) tiles += (q.e - q.s + 1 + (int64_t)kOmega * P->sigma - 1) / ((int64_t)kOmega * P->sigma);
    tiles_per_unit = 4 * std::min<int64_t>(kTileUnitChunks, std::max<int64_t>(1, tiles / (4 * 1024)));
  }
  int64_t uoff = 0, cooff = 0, toff = 0, tpoff = 0;
  std::vector<int32_t> vstart;  // kernel 1: first row of each unit
  for (auto &q : P->parts) {
    const int64_t rows = (int64_t)q.row_hi - q.row_lo;
    const int64_t nzq = q.e - q.s + 1;
    q.unit_off = uoff;
    q.carry_off = 2 * uoff;
    q.cslot1 = -1;
    if (P->kernel == SPX_KERNEL_VECTOR) {
      const size_t first = vstart.size();
      if (P->parts.size() == 1 && q.row_lo == 0 && q.row_hi == m && q.s == 0 && q.e + 1 == nnzr && !vwhole.empty())
        vstart = vwhole;  // cut before the upload ended
      else
        vec_units(rp, q.row_lo, q.row_hi, q.s, q.e + 1, P->lanes, nnzr, vstart);
      q.units = (int64_t)(vstart.size() - first);
    } else if (P->kernel == SPX_KERNEL_MERGE) {
      const int64_t L = (rows - q.eflag) + nzq;
      q.units = (L + tw - 1) / tw;
      q.coord_off = cooff;
      cooff += q.units + 1;
    } else {
      const int64_t w = (int64_t)kOmega * P->sigma;
      q.ntiles = (nzq + w - 1) / w;
      q.tile_off = toff;
      q.tp_off = tpoff;
      toff += q.ntiles;
      tpoff += q.ntiles + 1;
      q.units = (q.ntiles + tiles_per_unit - 1) / tiles_per_unit;
    }
    uoff += q.units;
  }
  P->nunits = uoff;
  P->ncarry = 2 * uoff;
  P->ncoords = cooff;
  P->ntiles = toff;
  P->ntp = tpoff;
  HIP_TRY(hipMemcpy(P->d_parts, P->parts.data(), sizeof(Part) * P->parts.size(), hipMemcpyHostToDevice));

  if (P->kernel == SPX_KERNEL_VECTOR) {
    // rows of each unit and the partition's nonzero range; the window fields
    // first carry the unit's nonzero range
    std::vector<VUnit> vu(std::max<int64_t>(1, P->nunits));
    for (const Part &q : P->parts)
      for (int64_t u = 0; u < q.units; u++) {
        VUnit &v = vu[q.unit_off + u];
        v.r_lo = vstart[(size_t)(q.unit_off + u)];
        v.r_hi = u + 1 < q.units ? vstart[(size_t)(q.unit_off + u + 1)] : q.row_hi;
        v.s = (int32_t)q.s;
        v.e1 = (int32_t)(q.e + 1);
        v.cmin = v.z0 = std::max<int32_t>(rp[v.r_lo], v.s);
        v.cwin = v.z1 = std::min<int32_t>(rp[v.r_hi], v.e1);
      }
    P->vnnz.resize((size_t)P->nunits);
    for (int64_t u = 0; u < P->nunits; u++) P->vnnz[(size_t)u] = (int64_t)vu[(size_t)u].z1 - vu[(size_t)u].z0;
    if ((rc = P->upload(&P->d_vunits, vu.data(), P->nunits))) return rc;
    HIP_TRY(launch_unit_window(P->d_col, reinterpret_cast<int32_t *>(P->d_vunits), 8, 4, P->nunits, 0));
  }
  if (P->kernel == SPX_KERNEL_TILE) {
    // units of tiles_per_unit tiles; the window fields first carry the
    // unit's nonzero range in the CSR arrays
    std::vector<TUnit> tu(std::max<int64_t>(1, P->nunits));
    const int64_t w = (int64_t)kOmega * P->sigma;
    for (const Part &q : P->parts) {
      const int64_t nzq = q.e - q.s + 1;
      for (int64_t u = 0; u < q.units; u++) {
        TUnit &t = tu[q.unit_off + u];
        const int64_t lt0 = u * tiles_per_unit;
        t.gt0 = (int32_t)(q.tile_off + lt0);
        t.nt = (int32_t)std::min<int64_t>(tiles_per_unit, q.ntiles - lt0);
        t.lt0 = (int32_t)lt0;
        t.tp0 = (int32_t)(q.tp_off + lt0);
        t.nz = (int32_t)nzq;
        t.row_lo = q.row_lo;
        t.s = (int32_t)q.s;
        t.gu = (int32_t)(q.unit_off + u);
        t.cmin = (int32_t)(q.s + lt0 * w);
        t.cwin = (int32_t)(q.s + std::min<int64_t>((lt0 + t.nt) * w, nzq));
        t.flags = (q.sflag ? kTuSplitHead : 0) | (q.eflag ? kTuSplitTail : 0);
        t.last_lr = q.row_hi - q.row_lo - 1;
      }
    }
    if ((rc = P->upload(&P->d_tunits, tu.data(), P->nunits))) return rc;
    HIP_TRY(launch_unit_window(P->d_col, reinterpret_cast<int32_t *>(P->d_tunits), 12, 8, P->nunits, 0));
  }

  if ((rc = P->alloc(&P->d_crow, P->ncarry))) return rc;
  if ((rc = P->alloc(&P->d_cval, P->ncarry))) return rc;
  HIP_TRY(launch_fill_i32(P->d_crow, -1, P->ncarry, 0));
  HIP_TRY(hipMemset(P->d_cval, 0, sizeof(double) * std::max<int64_t>(1, P->ncarry)));

  // ---- regular payload ----
  if (P->kernel == SPX_KERNEL_TILE) {
    // the CSR arrays become the tile layout (64 lanes x sigma per tile)
    const int64_t tot = P->ntiles * kOmega * P->sigma;
    int32_t *csr_col = P->d_col;
    double *csr_val = P->d_val;
    P->d_col = nullptr;
    P->d_val = nullptr;
    rc = P->alloc(&P->d_tp, P->ntp);
    if (!rc) rc = P->alloc(&P->d_tf, (P->ntiles + 1) * kOmega);  // one tile of padding: a lane reads its right neighbour's word
    if (!rc) rc = hipMemset(P->d_tf + P->ntiles * kOmega, 0, sizeof(uint32_t) * kOmega) == hipSuccess ? SPX_OK : SPX_EDEV;
    if (!rc) rc = P->alloc(&P->d_col, tot);
    if (!rc) rc = P->precision == SPX_PREC_F32 ? P->alloc(&P->d_val32, tot) : P->alloc(&P->d_val, tot);
    hipError_t e = hipSuccess;
    if (!rc && nnzr > 0)
      e = launch_tile_convert(P->d_rowptr, csr_col, csr_val, P->parts.data(), P->d_parts, (int)P->parts.size(),
                              P->sigma, P->d_tp, P->d_tf, P->d_col, P->d_val, P->d_val32, 0);
    if (!rc && e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(csr_col);
    (void)hipFree(csr_val);
    P->bytes -= (int64_t)(nnzr + 2) * (sizeof(int32_t) + sizeof(double));
    if (rc) return rc;
    HIP_TRY(e);
    // empty rows are written by their own small kernel (the long rows, empty
    // here, by the long-row fold)
    std::vector<int32_t> empty;
    size_t lj = 0;
    for (int i = 0; i < m; i++) {
      const bool is_long = lj < long_rows.size() && long_rows[lj] == i;
      if (is_long) lj++;
      if (rp[i] == rp[i + 1] && !is_long) empty.push_back(i);
    }
    P->nempty = (int64_t)empty.size();
    if ((rc = P->upload(&P->d_empty, empty.data(), P->nempty))) return rc;
  } else if (P->kernel == SPX_KERNEL_MERGE) {
    if ((rc = P->alloc(&P->d_coords, P->ncoords))) return rc;
    HIP_TRY(launch_merge_coords(P->d_rowptr, P->d_parts, (int)P->parts.size(), P->ncoords, tw, P->d_coords, 0));
    if ((rc = P->alloc(&P->d_munits, P->nunits))) return rc;
    for (const Part &q : P->parts)
      HIP_TRY(launch_merge_units(P->d_coords + q.coord_off, q.units, P->d_munits + q.unit_off, 0));
    HIP_TRY(launch_unit_window(P->d_col, reinterpret_cast<int32_t *>(P->d_munits), 8, 4, P->nunits, 0));
  }

  // ---- launch layout: partitions -> streams ----
  int nstreams = 1;
  if (P->mode == SPX_MODE_QUEUE) {
    nstreams = std::max(1, std::min(P->queues, (int)P->parts.size()));
    if ((rc = P->alloc(&P->d_queue, (int64_t)nstreams * kXcd))) return rc;
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int np2 = (int)P->parts.size();
    for (int c = 0; c < nstreams; c++) {
      // contiguous task quota per stream (SB/src/dspmv_mgpu_v2.cu:124-126)
      int t0 = (int)((int64_t)np2 * c / nstreams), t1 = (int)((int64_t)np2 * (c + 1) / nstreams);
      if (t1 <= t0) continue;
      int64_t ub = P->parts[t0].unit_off;
      int64_t ue = P->parts[t1 - 1].unit_off + P->parts[t1 - 1].units;
      if (ue <= ub) continue;
      // persistent workgroups: a multiple of the XCD count, ~8 per CU overall
      int64_t want = std::max<int64_t>(kXcd, (int64_t)cus * 8 / nstreams);
      int grid = (int)std::min<int64_t>(direct_grid(ue - ub), direct_grid(want));
      P->launches.push_back({c, ub, ue - ub, grid});
    }
  } else {
    // one launch per partition, partitions round-robin over at most 8 streams
    nstreams = std::min<int>((int)P->parts.size(), 8);
    for (size_t pi = 0; pi < P->parts.size(); pi++) {
      const Part &q = P->parts[pi];
      if (q.units == 0) continue;
      P->launches.push_back({(int)(pi % nstreams), q.unit_off, q.units, direct_grid(q.units)});
    }
    if (P->kernel == SPX_KERNEL_VECTOR && !P->launches.empty()) {
      // each launch's units split over the XCDs by nonzeros (plus a fixed
      // cost per unit): contiguous ranges, so neighbouring rows share an L2
      std::vector<int32_t> xr;
      for (auto &L : P->launches) {
        std::vector<double> w((size_t)L.ucnt + 1, 0.0);
        for (int64_t u = 0; u < L.ucnt; u++) w[(size_t)u + 1] = w[(size_t)u] + (double)P->vnnz[(size_t)(L.ubeg + u)] + 64.0;
        xr.push_back(0);
        for (int x = 1; x < kXcd; x++)
          xr.push_back(std::max<int32_t>(xr.back(), (int32_t)(std::lower_bound(w.begin(), w.end() - 1, w[(size_t)L.ucnt] * x / kXcd) - w.begin())));
        xr.push_back((int32_t)L.ucnt);
      }
      if ((rc = P->upload(&P->d_vxr, xr.data(), (int64_t)xr.size()))) return rc;
    }
    P->vnnz.clear();
    P->vnnz.shrink_to_fit();
  }
  if (nstreams > 1) {
    P->streams.resize(nstreams);
    P->done.resize(nstreams);
    for (int c = 0; c < nstreams; c++) {
      HIP_TRY(hipStreamCreateWithFlags(&P->streams[c], hipStreamNonBlocking));
      HIP_TRY(hipEventCreateWithFlags(&P->done[c], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreateWithFlags(&P->fork, hipEventDisableTiming));
  }

  HIP_TRY(hipDeviceSynchronize());
  return SPX_OK;
}

extern "C" {

int spx_plan_create(spx_plan_t *out, const spx_plan_desc *d) {
  if (!out || !d) return fail(SPX_EARG, "null argument");
  *out = nullptr;
  std::unique_ptr<spx_plan> P(new spx_plan());
  const auto t0 = std::chrono::steady_clock::now();
  int rc = plan_build(P.get(), d);
  P->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (rc != SPX_OK) return rc;
  *out = P.release();
  return SPX_OK;
}

int spx_plan_destroy(spx_plan_t p) {
  delete p;
  return SPX_OK;
}

}  // extern "C"

template <typename V>
static hipError_t launch_main(spx_plan *P, const KArgs<V> &base, const spx_plan::Launch &L, hipStream_t st) {
  KArgs<V> a = base;
  a.unit_begin = L.ubeg;
  a.unit_count = L.ucnt;
  a.queue = nullptr;
  if (P->d_vxr) a.vxr = P->d_vxr + (&L - P->launches.data()) * (kXcd + 1);
  if (P->mode == SPX_MODE_QUEUE) {
    a.queue = P->d_queue + (int64_t)L.stream * kXcd;
    hipError_t e = launch_reset(a.queue, kXcd, st);
    if (e != hipSuccess) return e;
  }
  if constexpr (std::is_same<V, double>::value) {
    if (P->kernel == SPX_KERNEL_VECTOR) return launch_vector_f64(a, P->lanes, L.grid, st);
    if (P->kernel == SPX_KERNEL_MERGE) return launch_merge_f64(a, P->items, L.grid, st);
    return launch_tile_f64(a, L.grid, st);
  } else {
    if (P->kernel == SPX_KERNEL_VECTOR) return launch_vector_f32(a, P->lanes, L.grid, st);
    if (P->kernel == SPX_KERNEL_MERGE) return launch_merge_f32(a, P->items, L.grid, st);
    return launch_tile_f32(a, L.grid, st);
  }
}

template <typename V>
static hipError_t run_regular(spx_plan *P, const KArgs<V> &a, hipStream_t st) {
  hipError_t e = hipSuccess;
  if (P->streams.empty()) {
    for (auto &L : P->launches)
      if ((e = launch_main(P, a, L, st)) != hipSuccess) return e;
    return e;
  }
  if ((e = hipEventRecord(P->fork, st)) != hipSuccess) return e;
  std::vector<char> used(P->streams.size(), 0);
  for (auto &L : P->launches) {
    hipStream_t s = P->streams[L.stream];
    if (!used[L.stream]) {
      if ((e = hipStreamWaitEvent(s, P->fork, 0)) != hipSuccess) return e;
      used[L.stream] = 1;
    }
    if ((e = launch_main(P, a, L, s)) != hipSuccess) return e;
  }
  for (size_t c = 0; c < P->streams.size(); c++) {
    if (!used[c]) continue;
    if ((e = hipEventRecord(P->done[c], P->streams[c])) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(st, P->done[c], 0)) != hipSuccess) return e;
  }
  return e;
}

template <typename V>
static KArgs<V> make_args(spx_plan *P, const V *val, const double *x, double alpha, double beta, double *y) {
  KArgs<V> a = {};
  a.parts = P->d_parts;
  a.units = P->d_units;
  a.rowptr = P->d_rowptr;
  a.col = P->d_col;
  a.val = val;
  a.x = x;
  a.alpha = alpha;
  a.beta = beta;
  a.y = y;
  a.crow = P->d_crow;
  a.cval = P->d_cval;
  a.vunits = P->d_vunits;
  a.munits = P->d_munits;
  a.tunits = P->d_tunits;
  a.tile_ptr = P->d_tp;
  a.tile_flags = P->d_tf;
  a.sigma = P->sigma;
  return a;
}

// The segmented engine over one class of units: SEG_ALL (one kernel; the
// one with the LDS window if any unit stages), SEG_STAGED or SEG_UNSTAGED.
enum { SEG_ALL = 0, SEG_STAGED = 1, SEG_UNSTAGED = 2 };
static hipError_t launch_seg(spx_plan *P, int cls, const double *x, double alpha, double beta, double *y,
                             hipStream_t st) {
  const int64_t cnt = cls == SEG_ALL ? P->nsunits : cls == SEG_STAGED ? P->nsstaged : P->nsunits - P->nsstaged;
  if (cnt <= 0) return hipSuccess;
  const int32_t *xr = P->d_sxr + cls * (kXcd + 1);
  const bool stage = cls == SEG_STAGED || (cls == SEG_ALL && P->sstage);
  const bool wide_out = P->nbands > 0 && !stage;  // column-band slices: one or two elements each
  if (P->precision == SPX_PREC_F32) {
    SegArgs<float, double> a = {P->d_sunits, cnt, xr, P->d_sstart, P->d_scol12, P->d_scol32, P->d_sval32,
                                x, alpha, beta, y, P->d_partial};
    return launch_seg_f32(a, stage, wide_out, st);
  }
  SegArgs<double, double> a = {P->d_sunits, cnt, xr, P->d_sstart, P->d_scol12, P->d_scol32, P->d_sval,
                               x, alpha, beta, y, P->d_partial};
  return launch_seg_f64(a, stage, wide_out, st);
}

// Enqueue one SpMV.  t0/t1 (optional) bracket the SpMV kernels proper (the
// regular partitions and the long-row sweep, which overlap); x splitting,
// the fix-up and the long-row reduction fall outside.  `dry` runs only the
// regular kernels (used once at plan time to record the carry runs).
static int execute_impl(spx_plan *P, double alpha, const double *x, double beta, double *y, hipStream_t st,
                        hipEvent_t t0, hipEvent_t t1, bool dry) {
  if (!P) return fail(SPX_EARG, "null plan");
  if (P->m == 0) return SPX_OK;
  if (!y || (P->n > 0 && !x)) return fail(SPX_EARG, "null x or y");
  if (P->nnz == 0) {  // no nonzeros (x may be empty): y = beta * y
    HIP_TRY(launch_scale_y(beta, y, P->m, st));
    return SPX_OK;
  }
  const bool f32 = P->precision == SPX_PREC_F32;
  if (P->kernel == SPX_KERNEL_SEG) {
    // rows of A without nonzeros; the long rows are written by the reduction
    if (P->nempty > 0) HIP_TRY(launch_empty_rows(P->d_empty, P->nempty, beta, y, st));
    if (t0) HIP_TRY(hipEventRecord(t0, st));
    if (P->seg_split) {  // the unstaged units on the second stream, concurrently
      HIP_TRY(hipEventRecord(P->lfork, st));
      HIP_TRY(hipStreamWaitEvent(P->lstream, P->lfork, 0));
      HIP_TRY(launch_seg(P, SEG_STAGED, x, alpha, beta, y, st));
      HIP_TRY(launch_seg(P, SEG_UNSTAGED, x, alpha, beta, y, P->lstream));
      HIP_TRY(hipEventRecord(P->ldone, P->lstream));
      HIP_TRY(hipStreamWaitEvent(st, P->ldone, 0));
    } else {
      HIP_TRY(launch_seg(P, SEG_ALL, x, alpha, beta, y, st));
    }
    // column bands: the regular rows' slices, added per row in band order
    if (P->nbands > 0) HIP_TRY(launch_band_reduce(P->d_bmask, P->d_bbase, P->nbands, P->d_partial, P->nlseg + P->nbslots, P->m, alpha,
                                        beta, y, st));
    if (t1) HIP_TRY(hipEventRecord(t1, st));
    if (P->nlong > 0)
      HIP_TRY(launch_long_reduce(P->d_partial, P->nlong, P->d_lsp, P->d_lseg, P->d_lrows, alpha, beta, true, y, st));
    return SPX_OK;
  }
  if (t0) HIP_TRY(hipEventRecord(t0, st));
  // The long rows' slices run first, on the same stream: the regular kernel
  // and the segmented engine each fill the GPU with persistent workgroups,
  // and sharing it between them measured slower than one after the other.
  const bool longpath = P->nlong > 0 && !dry;
  if (longpath) HIP_TRY(launch_seg(P, SEG_ALL, x, alpha, beta, y, st));
  if (P->kernel == SPX_KERNEL_TILE) HIP_TRY(launch_empty_rows(P->d_empty, P->nempty, beta, y, st));
  const bool regular = !(P->nnz_reg == 0 && P->kernel == SPX_KERNEL_TILE);
  if (regular) {
    hipError_t e;
    if (f32)
      e = run_regular(P, make_args<float>(P, P->d_val32, x, alpha, beta, y), st);
    else
      e = run_regular(P, make_args<double>(P, P->d_val, x, alpha, beta, y), st);
    HIP_TRY(e);
  }
  if (t1) HIP_TRY(hipEventRecord(t1, st));
  if (dry) return SPX_OK;
  // kernels 1 and 2 write the long rows as empty rows (beta * y), the fold
  // adds alpha * sum; kernel 3 leaves them to the fold
  HIP_TRY(launch_finish(P->d_partial, longpath ? P->nlong : 0, P->d_lsp, P->d_lseg, P->d_lrows,
                        P->kernel == SPX_KERNEL_TILE, P->d_crow, P->d_cval, P->ncarry, alpha, beta, y, st));
  return SPX_OK;
}

extern "C" {

int spx_plan_execute(spx_plan_t p, double alpha, const double *x, double beta, double *y, void *stream) {
  return execute_impl(p, alpha, x, beta, y, (hipStream_t)stream, nullptr, nullptr, false);
}

int spx_plan_execute_multi(spx_plan_t p, int k, double alpha, const double *X, int64_t ldx, double beta, double *Y,
                           int64_t ldy, void *stream) {
  if (!p || k < 0 || ldx < p->n || ldy < p->m) return fail(SPX_EARG, "bad multi-RHS arguments");
  if (p->m == 0 || k == 0) return SPX_OK;
  if (!X || !Y) return fail(SPX_EARG, "null X or Y");
  if (p->fused_spmm && p->kernel == SPX_KERNEL_SEG && p->precision == SPX_PREC_F64 && p->nbands == 0) {
    // fused passes of up to kSpmmPass columns: A is read once per pass
    hipStream_t st = (hipStream_t)stream;
    const int K = std::min(k, kSpmmPass);
    if (p->nlong > 0 && p->npartk < p->nlseg * K) {
      HIP_TRY(hipDeviceSynchronize());  // earlier passes may still read the scratch being replaced
      if (p->d_partk) HIP_TRY(hipFree(p->d_partk));
      p->d_partk = nullptr;
      p->npartk = 0;
      int rc = p->alloc(&p->d_partk, p->nlseg * K);
      if (rc) return rc;
      p->npartk = p->nlseg * K;
    }
    for (int j0 = 0; j0 < k; j0 += kSpmmPass) {
      SpmmArgs a = {};
      a.units = p->d_sunits;
      a.nunits = p->nsunits;
      a.xr = p->d_sxr;
      a.start = p->d_sstart;
      a.col12 = p->d_scol12;
      a.col32 = p->d_scol32;
      a.val = p->d_sval;
      a.X = X + (size_t)j0 * ldx;
      a.ldx = ldx;
      a.kk = std::min(kSpmmPass, k - j0);
      a.alpha = alpha;
      a.beta = beta;
      a.Y = Y + (size_t)j0 * ldy;
      a.ldy = ldy;
      a.empty = p->d_empty;
      a.nempty = p->nempty;
      a.partial = p->d_partk;
      a.pstride = p->nlseg;
      a.nlong = p->nlong;
      a.lsp = p->d_lsp;
      a.lseg = p->d_lseg;
      a.lrows = p->d_lrows;
      HIP_TRY(launch_spmm(a, st));
    }
    return SPX_OK;
  }
  for (int j = 0; j < k; j++) {
    int rc = execute_impl(p, alpha, X + (size_t)j * ldx, beta, Y + (size_t)j * ldy, (hipStream_t)stream, nullptr,
                          nullptr, false);
    if (rc) return rc;
  }
  return SPX_OK;
}

int spx_plan_info(spx_plan_t p, int64_t *out, int cap) {
  if (!p || !out) return 0;
  int64_t v[] = {p->m,      p->n,        p->nnz,       p->kernel, p->mode,   (int64_t)p->parts.size(),
                 p->sigma,  p->lanes,    p->items,
                 p->kernel == SPX_KERNEL_SEG ? p->nsunits : p->nunits, p->ncarry, p->ntiles,
                 p->bytes,  p->precision, p->queues,   p->nlong,  p->long_nnz, p->nblocks,
                 0,  p->long_th, p->nstaged_units, p->seg_split ? 1 : 0,
                 (int64_t)(p->upload_ms * 1000.0), (int64_t)(p->build_ms * 1000.0), p->nbands};
  int c = std::min<int>(cap, (int)(sizeof(v) / sizeof(v[0])));
  for (int i = 0; i < c; i++) out[i] = v[i];
  return c;
}

int64_t spx_plan_export(spx_plan_t p, int what, void *dst, int64_t cap) {
  if (!p) return -1;
  const void *src = nullptr;
  int64_t bytes = 0;
  std::vector<int64_t> tmp;
  switch (what) {
    case 0:
      for (auto &q : p->parts) {
        tmp.push_back(q.s); tmp.push_back(q.e); tmp.push_back(q.r0);
        tmp.push_back(q.r1); tmp.push_back(q.sflag); tmp.push_back(q.eflag);
      }
      bytes = (int64_t)tmp.size() * 8;
      if (dst) std::memcpy(dst, tmp.data(), (size_t)std::min(bytes, cap));
      return bytes;
    case 1: src = p->d_tp; bytes = p->ntp * 4; break;
    case 2: src = p->d_tf; bytes = p->ntiles * kOmega * 4; break;
    case 3: src = p->kernel == SPX_KERNEL_TILE ? p->d_col : nullptr; bytes = p->ntiles * kOmega * p->sigma * 4; break;
    case 4: src = p->kernel == SPX_KERNEL_TILE ? p->d_val : nullptr; bytes = p->ntiles * kOmega * p->sigma * 8; break;
    case 5: {
      if (!p->d_coords) return 0;
      std::vector<Coord> c((size_t)p->ncoords);
      if (dst && p->ncoords &&
          hipMemcpy(c.data(), p->d_coords, sizeof(Coord) * p->ncoords, hipMemcpyDeviceToHost) != hipSuccess)
        return -1;
      for (auto &x : c) { tmp.push_back(x.row); tmp.push_back(x.nz); }
      bytes = p->ncoords * 16;
      if (dst) std::memcpy(dst, tmp.data(), (size_t)std::min(bytes, cap));
      return bytes;
    }
    default: return -1;
  }
  if (!src) return 0;
  if (dst && bytes > 0 && hipMemcpy(dst, src, (size_t)std::min(bytes, cap), hipMemcpyDeviceToHost) != hipSuccess)
    return -1;
  return bytes;
}

int spx_plan_time_each(spx_plan_t p, double alpha, const double *x, double beta, double *y, int warmup, int reps,
                       int64_t flush_bytes, double *exec_ms, double *kernel_ms, double *wall_ms) {
  const bool events = exec_ms != nullptr;
  if (!p || reps <= 0 || events != (kernel_ms != nullptr) || (!events && !wall_ms))
    return fail(SPX_EARG, "bad timing arguments");
  // Every execute gets its own event pairs (unless only the wall clock is
  // asked for), so the timed executes run back-to-back on one stream with a
  // single host wait at the end.
  hipStream_t st = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  std::vector<hipEvent_t> ev(events ? 4 * (size_t)reps : 0, nullptr);
  int rc = SPX_OK;
  for (auto &e : ev)
    if (rc == SPX_OK && hipEventCreate(&e) != hipSuccess) rc = fail(SPX_EDEV, "hipEventCreate failed");
  void *flush = nullptr;
  const int64_t fb = flush_bytes < 0 ? -flush_bytes : flush_bytes;
  if (rc == SPX_OK && fb > 0) {
    if (dev_malloc(&flush, (size_t)fb + 256) != hipSuccess)
      rc = fail(SPX_EMEM, "hipMalloc of the cache-flush buffer failed");
    else if (hipMemset(flush, 0, (size_t)fb + 256) != hipSuccess)
      rc = fail(SPX_EDEV, "hipMemset of the cache-flush buffer failed");
  }
  unsigned *sink = flush ? reinterpret_cast<unsigned *>(static_cast<char *>(flush) + fb) : nullptr;
  for (int i = 0; i < warmup && rc == SPX_OK; i++) rc = execute_impl(p, alpha, x, beta, y, st, nullptr, nullptr, false);
  if (rc == SPX_OK && hipStreamSynchronize(st) != hipSuccess) rc = fail(SPX_EDEV, "stream synchronize failed");
  const auto w0 = std::chrono::steady_clock::now();
  for (int i = 0; i < reps && rc == SPX_OK; i++) {
    if (flush && flush_bytes > 0) (void)launch_flush_read(flush, fb, sink, st);
    if (flush && flush_bytes < 0) (void)hipMemsetAsync(flush, i & 0xff, (size_t)fb, st);
    if (!events) {
      rc = execute_impl(p, alpha, x, beta, y, st, nullptr, nullptr, false);
      continue;
    }
    (void)hipEventRecord(ev[4 * i], st);
    rc = execute_impl(p, alpha, x, beta, y, st, ev[4 * i + 1], ev[4 * i + 2], false);
    (void)hipEventRecord(ev[4 * i + 3], st);
  }
  if (hipStreamSynchronize(st) != hipSuccess && rc == SPX_OK) rc = fail(SPX_EDEV, "stream synchronize failed");
  const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
  if (rc == SPX_OK) {
    for (int i = 0; i < reps && events; i++) {
      float t = 0, tk = 0;
      (void)hipEventElapsedTime(&t, ev[4 * i], ev[4 * i + 3]);
      (void)hipEventElapsedTime(&tk, ev[4 * i + 1], ev[4 * i + 2]);
      exec_ms[i] = t;
      kernel_ms[i] = tk;
    }
    if (wall_ms) *wall_ms = wall / reps;
  }
  if (flush) (void)hipFree(flush);
  for (auto e : ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(st);
  return rc;
}

int spx_plan_time(spx_plan_t p, double alpha, const double *x, double beta, double *y, int warmup, int reps,
                  int64_t flush_bytes, double *out_ms) {
  if (!p || reps <= 0 || !out_ms) return fail(SPX_EARG, "bad timing arguments");
  std::vector<double> e((size_t)reps), k((size_t)reps);
  double wall = 0.0;
  const int rc = spx_plan_time_each(p, alpha, x, beta, y, warmup, reps, flush_bytes, e.data(), k.data(), &wall);
  if (rc != SPX_OK) return rc;
  out_ms[0] = std::accumulate(e.begin(), e.end(), 0.0) / reps;
  out_ms[1] = std::accumulate(k.begin(), k.end(), 0.0) / reps;
  out_ms[2] = wall;
  return SPX_OK;
}

// ---- device memory helpers -------------------------------------------------------

void *spx_malloc(int64_t bytes) {
  void *p = nullptr;
  if (dev_malloc(&p, (size_t)(bytes > 0 ? bytes : 1)) != hipSuccess) {
    g_last_error = "hipMalloc failed";
    return nullptr;
  }
  return p;
}
int spx_free(void *p) {
  HIP_TRY(hipFree(p));
  return SPX_OK;
}
int spx_memcpy_h2d(void *dst, const void *src, int64_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  return SPX_OK;
}
int spx_memcpy_d2h(void *dst, const void *src, int64_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return SPX_OK;
}
int spx_memcpy_d2d(void *dst, const void *src, int64_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice));
  return SPX_OK;
}
int spx_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return SPX_OK;
}
int spx_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

}  // extern "C"

// ---- batched form: block-diagonal packing ------------------------------------------------

extern "C" int spx_batch_create(spx_plan_t *out, int count, const spx_batch_item *items, const spx_plan_desc *tmpl,
                                int64_t *x_off, int64_t *y_off) {
  if (!out || count < 0 || (count > 0 && !items) || !tmpl) return fail(SPX_EARG, "bad batch arguments");
  int64_t M = 0, N = 0, NNZ = 0;
  for (int i = 0; i < count; i++) {
    const spx_batch_item &it = items[i];
    if (it.m < 0 || it.n < 0 || it.nnz < 0 || (it.m > 0 && !it.rowptr) || (it.nnz > 0 && (!it.col || !it.val)))
      return fail(SPX_EARG, "bad batch item");
    if (x_off) x_off[i] = N;
    if (y_off) y_off[i] = M;
    M += it.m;
    N += it.n;
    NNZ += it.nnz;
  }
  if (x_off) x_off[count] = N;
  if (y_off) y_off[count] = M;
  if (M >= INT32_MAX || N >= INT32_MAX) return fail(SPX_ERANGE, "batch too large: sum of m or n must be below 2^31");
  if (NNZ >= ((int64_t)1 << 31) - 1) return fail(SPX_ERANGE, "batch too large: total nnz must be below 2^31");
  std::vector<int64_t> rp((size_t)M + 1, 0);
  std::vector<int32_t> col((size_t)NNZ);
  std::vector<double> val((size_t)NNZ);
  int64_t r0 = 0, c0 = 0, z0 = 0;
  for (int i = 0; i < count; i++) {
    const spx_batch_item &it = items[i];
    // each member's row pointer runs monotonically from 0 to its nnz (a member
    // with no rows owns no nonzeros), so no member's nonzeros land in another
    if (it.m == 0 && it.nnz != 0) return fail(SPX_EARG, "batch item without rows has nonzeros");
    if (it.m > 0 && (it.rowptr[0] != 0 || it.rowptr[it.m] != it.nnz)) return fail(SPX_EARG, "bad batch row pointer");
    for (int r = 1; r <= it.m; r++) {
      if (it.rowptr[r] < it.rowptr[r - 1]) return fail(SPX_EARG, "batch row pointer not monotone");
      rp[r0 + r] = z0 + it.rowptr[r];
    }
    for (int64_t k = 0; k < it.nnz; k++) {
      if (it.col[k] < 0 || it.col[k] >= it.n) return fail(SPX_EARG, "batch column index out of range");
      col[z0 + k] = (int32_t)(it.col[k] + c0);
    }
    std::memcpy(val.data() + z0, it.val, sizeof(double) * (size_t)it.nnz);
    r0 += it.m;
    c0 += it.n;
    z0 += it.nnz;
  }
  spx_plan_desc d = *tmpl;
  d.m = (int)M;
  d.n = (int)N;
  d.nnz = NNZ;
  d.rowptr64 = rp.data();
  d.rowptr32 = nullptr;
  d.col = col.data();
  d.val = val.data();
  d.flags &= ~SPX_FLAG_DEVICE_PTRS;
  return spx_plan_create(out, &d);
}

// ---- HIP graphs ----------------------------------------------------------------------------

struct spx_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~spx_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

extern "C" int spx_graph_capture(spx_plan_t p, double alpha, const double *x, double beta, double *y,
                                 spx_graph_t *out) {
  if (!p || !out) return fail(SPX_EARG, "null argument");
  *out = nullptr;
  hipStream_t st = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  std::unique_ptr<spx_graph> G(new spx_graph());
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  int rc = SPX_OK;
  if (e == hipSuccess) {
    rc = execute_impl(p, alpha, x, beta, y, st, nullptr, nullptr, false);
    hipError_t e2 = hipStreamEndCapture(st, &G->graph);
    if (rc == SPX_OK && e2 != hipSuccess) e = e2;
  }
  if (rc == SPX_OK && e == hipSuccess) e = hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0);
  (void)hipStreamDestroy(st);
  if (rc != SPX_OK) return rc;
  HIP_TRY(e);
  *out = G.release();
  return SPX_OK;
}

extern "C" int spx_graph_launch(spx_graph_t g, void *stream) {
  if (!g) return fail(SPX_EARG, "null graph");
  HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return SPX_OK;
}

extern "C" int spx_graph_destroy(spx_graph_t g) {
  delete g;
  return SPX_OK;
}

extern "C" int spx_graph_time(spx_graph_t g, int warmup, int reps, double *out_ms) {
  if (!g || reps <= 0 || !out_ms) return fail(SPX_EARG, "bad timing arguments");
  hipStream_t st = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  for (int i = 0; i < warmup && e == hipSuccess; i++) e = hipGraphLaunch(g->exec, st);
  if (e == hipSuccess) e = hipEventRecord(a, st);
  for (int i = 0; i < reps && e == hipSuccess; i++) e = hipGraphLaunch(g->exec, st);
  if (e == hipSuccess) e = hipEventRecord(b, st);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  (void)hipStreamDestroy(st);
  HIP_TRY(e);
  out_ms[0] = ms / reps;
  return SPX_OK;
}
