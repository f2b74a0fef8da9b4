// spx_internal.h -- types shared by the gfx950 kernels and the host plan code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace spx {

constexpr int kBlock = 256;   // threads per workgroup (4 wave64s)
constexpr int kWave = 64;     // CDNA wavefront width
constexpr int kOmega = 64;    // lanes per CSR5-style tile
constexpr int kXcd = 8;       // accelerator complex dies (one L2 each)
constexpr int kColBlock = 2560;   // long-row path: column block width (x staged in LDS)
// Long-row slices are cut into units per column block of at least this many
// nonzeros, so a block is one unit (its x window staged once) unless it is
// denser than that.
constexpr int kLongUnit = 32768;
constexpr int kMaxLong = 1024;    // long-row path: most long rows per plan

// Workgroups for a direct launch over `count` units: a multiple of kXcd so
// every XCD receives the same contiguous share.
__host__ __device__ inline int direct_grid(int64_t count) {
  return (int)(((count + kXcd - 1) / kXcd) * kXcd);
}

// One nnz-contiguous partition of the matrix (a stream's share in the v1
// mode, a task in the v2 mode, a row block in the baseline mode).
// [s, e] is the nonzero range (e inclusive, as SB/src/dspmv_mgpu_v1.cu:62-69),
// r0/r1 the rows holding nonzeros s and e, sflag/eflag whether those rows are
// split at the partition boundary (v1:77, v1:89).  The partition owns rows
// [row_lo, row_hi): r0..r1 plus the empty rows up to the next partition.
struct Part {
  int64_t s, e;
  int32_t r0, r1, sflag, eflag;
  int32_t row_lo, row_hi;
  int64_t units;      // workgroup work items of this partition
  int64_t unit_off;   // first entry in the unit table
  int64_t carry_off;  // carry slot of the head split row; unit u's slots follow
                      // at carry_off + 1 + 2u (+1) (per tile for the tile kernel)
  int64_t cslot1;     // carry slot of the tail split row
  int64_t coord_off;  // merge path: unit boundary coordinates (units + 1)
  int64_t ntiles;     // tiles: count
  int64_t tile_off;   // tiles: first global tile index
  int64_t tp_off;     // tiles: first tile pointer (ntiles + 1 per partition)
};

// A unit of work: (partition, unit index within the partition).
struct Unit {
  int32_t part, local;
};

// Merge-path coordinate of a unit boundary: absolute row, absolute nonzero.
struct Coord {
  int32_t row, nz;
};


// Kernel 2 (merge path): one unit of the merge of row ends and nonzeros,
// from coordinate (row r0, nonzero j0) to (r1, j1).  Rows r0 .. r1-1 end in
// the unit; nonzeros j0 .. j1-1 are consumed by it.
struct MUnit {
  int32_t r0, j0, r1, j1;
  int32_t cmin, cwin;  // x window staged in LDS (cwin = 0: gathers from x)
  int32_t pad0, pad1;
};

// Kernel 1 (CSR-vector): one unit of consecutive rows [r_lo, r_hi) inside a
// partition whose nonzeros are [s, e1).  A row cut by the partition's bounds
// leaves its piece in carry slot 2u (the unit's first row) or 2u + 1 (its
// last row), u the unit's global index.
struct VUnit {
  int32_t r_lo, r_hi, s, e1;
  int32_t cmin, cwin;  // x window staged in LDS (cwin = 0: gathers from x)
  int32_t z0, z1;      // the unit's nonzeros (clamped to the partition)
};
// Kernel 1 stages a unit's products in LDS when it has at most
// kBlock * kVecItems nonzeros (each thread loads kVecItems of them).
constexpr int kVecItems = 8;

// Kernel 3 (CSR5-style tiles): tiles [lt0, lt0 + nt) of one partition, in
// chunks of 4 (one per wave).  Carry slots 2gu (the piece of the row open at
// the unit's start) and 2gu + 1 (the piece open at its end).
struct TUnit {
  int32_t gt0;      // global index of the first tile (tile storage, flags)
  int32_t nt;       // tiles
  int32_t lt0;      // the first tile's index within its partition
  int32_t tp0;      // tile pointer entry of the first tile
  int32_t nz;       // the partition's nonzeros
  int32_t row_lo;   // the partition's first row
  int32_t s;        // the partition's first nonzero
  int32_t gu;       // global unit index
  int32_t cmin, cwin;
  int32_t flags;    // kTuSplitHead: the partition's first row began earlier;
                    // kTuSplitTail: its last row continues after it
  int32_t last_lr;  // the partition's last local row
};
enum { kTuSplitHead = 1, kTuSplitTail = 2 };
constexpr int kTileUnitChunks = 4;  // chunks of 4 tiles per kernel-3 unit
constexpr int kTileOut = 512;       // row sums per chunk gathered in LDS before they are stored
// Kernel 1's units: consecutive rows of one partition, at most kVecUnitRows
// of them and at most kBlock * kVecItems nonzeros, except a longer row, which
// is a unit alone (the plan may cap both lower).
constexpr int kVecUnitRows = 1024;

// Carry slots hold pieces of rows split between units, in nonzero order.  A
// slot's row word is -1 when unused, else the row with two flag bits: the
// row's first piece (kSlotFirst) and its final piece (kSlotFinal).  Pieces
// between the two carry neither.  A row's pieces are consecutive among the
// used slots, so the final piece's fix-up walks back to the first.
constexpr int32_t kSlotFirst = 1 << 30;
constexpr uint32_t kSlotFinal = 1u << 31;
constexpr int32_t kSlotRow = kSlotFirst - 1;

// ---- segmented engine (seg.hip) ----------------------------------------------
constexpr int kSegStage = 3328;  // most x entries a unit stages in LDS (26 KiB)
static_assert(kSegStage <= 4096, "staged column offsets are 12-bit");
constexpr int kSegBlock = 256;   // threads per segmented-engine workgroup (4 wave64s)
constexpr int kSegOut = 520;     // sums per chunk gathered in LDS before they are stored
constexpr int kSegOutWide = kSegBlock * 8 + 8;  // every sum a chunk can close

// One workgroup unit of the segmented engine: padded elements [e0, e1)
// (e0 a multiple of 8) holding segments s0 .. s0 + ns - 1 in order.  A
// SEG_STAGE unit's columns lie in [cmin, cmin + cwin) and are stored as
// 12-bit offsets from cmin (col12: element e's offset in bits 12 (e & 7) ..
// 12 (e & 7) + 11 of the three 32-bit words 3 (e >> 3) .. 3 (e >> 3) + 2);
// any other unit's columns are stored in full in col32 from index cmin on.  Segment s0 + j writes output o0 + j: y
// row o0 + j, or partial slot o0 + j in SEG_PARTIAL units.  Bit (e & 7) of
// start[e >> 3] is set when element e begins a segment.
struct SegUnit {
  int32_t e0, e1, s0, ns, o0, cmin, cwin, flags;
};

enum { SEG_STAGE = 1, SEG_PARTIAL = 2 };  // SegUnit.flags

// Physical order of the segmented engine's column and value arrays.  From
// e0, a unit's elements form wave blocks of 512 (the last one shorter); in
// block B, lane l owns elements B + 8l .. B + 8l + 7, and with L the block's
// lanes (its elements rounded up to 8, over 8) element B + 8l + i sits at
// B + (i/4)*4L + 4l + i%4 in the 4-byte arrays (columns, FP32 values) and at
// B + (i/2)*2L + 2l + i%2 in the FP64 values.  Each 16-byte load instruction
// of a wave then reads one contiguous run of 64 x 16 bytes.
constexpr int kSegWaveBlock = 512;
__host__ __device__ inline int seg_lanes(int64_t B, int64_t e1) {
  const int64_t L = (e1 - B + 7) >> 3;
  return L > 64 ? 64 : (int)L;
}
__host__ __device__ inline int64_t seg_pos4(int64_t e, int64_t e0, int64_t e1) {
  const int64_t B = e0 + ((e - e0) & ~(int64_t)(kSegWaveBlock - 1));
  const int off = (int)(e - B), l = off >> 3, i = off & 7;
  return B + (int64_t)(i >> 2) * 4 * seg_lanes(B, e1) + 4 * l + (i & 3);
}
__host__ __device__ inline int64_t seg_pos8(int64_t e, int64_t e0, int64_t e1) {
  const int64_t B = e0 + ((e - e0) & ~(int64_t)(kSegWaveBlock - 1));
  const int off = (int)(e - B), l = off >> 3, i = off & 7;
  return B + (int64_t)(i >> 1) * 2 * seg_lanes(B, e1) + 2 * l + (i & 1);
}

template <typename V, typename X>
struct SegArgs {
  const SegUnit *units;
  int64_t nunits;
  const int32_t *xr;     // XCD x owns units [xr[x], xr[x + 1]) (kXcd + 1 entries)
  const uint8_t *start;  // segment-start bits, one per padded element
  const uint32_t *col12;  // staged units: column - cmin, 12 bits each, 8 per 3 words
  const int32_t *col32;   // unstaged units: columns from cmin on, lane-fastest (seg_pos4 - e0)
  const V *val;
  const X *x;
  double alpha, beta;
  double *y;
  double *partial;
};

// One multi-RHS pass (spmm.hip): columns of X and Y starting at the given
// pointers, kk <= kSpmmPass of them.
constexpr int kSpmmPass = 64;  // right-hand sides per fused SpMM pass (A read once per pass)

struct SpmmArgs {
  const SegUnit *units;
  int64_t nunits;
  const int32_t *xr;
  const uint8_t *start;  // segment-start bits, one byte per 8 elements
  const uint32_t *col12;  // as in SegArgs
  const int32_t *col32;
  const double *val;
  const double *X;  // the pass's first column
  int64_t ldx;
  int kk;  // columns in this pass, 1..kSpmmPass
  double alpha, beta;
  double *Y;  // the pass's first column
  int64_t ldy;
  const int32_t *empty;  // rows of A without nonzeros (long rows excluded)
  int64_t nempty;
  double *partial;  // slice o of column j at partial[j * pstride + o]
  int64_t pstride;
  int nlong;
  const int32_t *lsp, *lseg, *lrows;
};

template <typename V>
struct KArgs {
  const Part *parts;
  const Unit *units;
  int64_t unit_begin, unit_count;
  int *queue;  // non-null: persistent workgroups pop units from this counter
  const int32_t *rowptr;
  const int32_t *col;
  const V *val;
  const double *x;
  double alpha, beta;
  double *y;
  int32_t *crow;
  double *cval;
  const VUnit *vunits;       // kernel 1
  const int32_t *vxr;        // kernel 1 without the queue: XCD x walks units [vxr[x], vxr[x+1])
  const MUnit *munits;       // kernel 2
  const TUnit *tunits;       // kernel 3
  const uint32_t *tile_ptr;  // kernel 3
  const uint32_t *tile_flags;
  int sigma;
};

// Copies between pageable host memory and the device through a pinned double
// buffer (stage.cpp); small copies go straight through hipMemcpy.
hipError_t stage_h2d(void *dst, const void *src, size_t bytes);
hipError_t stage_d2h(void *dst, const void *src, size_t bytes);

// hipMalloc behind the allocation-failure test seam (spx_debug_fail_alloc).
hipError_t dev_malloc(void **p, size_t bytes);

// The thread's message for spx_last_error (plan.cpp).
void set_last_error(const std::string &msg);

// ---- launchers (kernels.hip) ---------------------------------------------
enum { PART_NNZ = 0, PART_TASKS = 1, PART_ROWS = 2 };

// Device partitioner: fills parts[0..P) (s, e, r0, r1, flags, row_lo/hi).
hipError_t launch_partition(const int32_t *rowptr, int m, int64_t nnz, int P,
                            int kind, Part *parts, hipStream_t st);
hipError_t launch_merge_coords(const int32_t *rowptr, const Part *parts,
                               int nparts, int64_t total_coords, int tw,
                               Coord *coords, hipStream_t st);
hipError_t launch_tile_convert(const int32_t *rowptr, const int32_t *col,
                               const double *val, const Part *hparts,
                               const Part *dparts, int nparts, int sigma,
                               uint32_t *tile_ptr, uint32_t *flags,
                               int32_t *colt, double *valt, float *valt32,
                               hipStream_t st);
hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t st);
hipError_t launch_to_f32(const double *v, float *o, int64_t n, hipStream_t st);
hipError_t launch_empty_rows(const int32_t *rows, int64_t nrows, double beta,
                             double *y, hipStream_t st);
// With S = sum of partial[lseg[j]], j in [lsp[r], lsp[r+1]): y[rows[r]] =
// alpha*S + beta*y[rows[r]] (`assign`), else y[rows[r]] += alpha*S.
hipError_t launch_long_reduce(const double *partial, int nlong, const int32_t *lsp,
                              const int32_t *lseg, const int32_t *rows, double alpha, double beta,
                              bool assign, double *y, hipStream_t st);
hipError_t launch_reset(int *q, int n, hipStream_t st);
hipError_t launch_find_empty(const int32_t *rowptr, int m, int32_t *flag,
                             hipStream_t st);

// SpMV main kernels.  grid = number of workgroups (persistent when a.queue).
hipError_t launch_vector_f64(const KArgs<double> &a, int lanes, int grid, hipStream_t st);
hipError_t launch_vector_f32(const KArgs<float> &a, int lanes, int grid, hipStream_t st);
hipError_t launch_merge_f64(const KArgs<double> &a, int items, int grid, hipStream_t st);
hipError_t launch_merge_f32(const KArgs<float> &a, int items, int grid, hipStream_t st);
hipError_t launch_tile_f64(const KArgs<double> &a, int grid, hipStream_t st);
hipError_t launch_tile_f32(const KArgs<float> &a, int grid, hipStream_t st);
// Fix-up of kernels 1-3: every slot holding a final piece sums its row's
// pieces (first .. final, in slot order) into y.
hipError_t launch_slot_fix(const int32_t *crow, const double *cval, int64_t nslots, double alpha, double beta,
                           double *y, hipStream_t st);
// The long-row fold (k_long_reduce) and the slot fix-up in one launch.
hipError_t launch_finish(const double *partial, int nlong, const int32_t *lsp, const int32_t *lseg,
                         const int32_t *rows, bool assign, const int32_t *crow, const double *cval, int64_t nslots,
                         double alpha, double beta, double *y, hipStream_t st);
// Per-unit x windows: for unit u, the int32 fields f[u*stride + off] and
// f[u*stride + off + 1] hold the unit's nonzero range [lo, hi) of `col` on
// entry and its staged window (cmin, cwin; cwin = 0 when wider than
// kSegStage or empty) on return.
hipError_t launch_unit_window(const int32_t *col, int32_t *f, int stride, int off, int64_t nunits, hipStream_t st);
// MUnit table of a partition from its coordinates (units + 1 entries).
hipError_t launch_merge_units(const Coord *coords, int64_t units, MUnit *out, hipStream_t st);

// Kernel 4 with column bands: y for every regular row from its band slices
// (seg.hip).
hipError_t launch_band_reduce(const uint32_t *mask, const int32_t *base, int nbands, const double *partial,
                              int64_t nslots, int m, double alpha, double beta, double *y, hipStream_t st);

// Multi-RHS SpMM pass on the segmented engine's layout (spmm.hip).
hipError_t launch_spmm(const SpmmArgs &a, hipStream_t st);

// Segmented engine: `stage` = some unit stages x in LDS (SEG_STAGE).
hipError_t launch_seg_f64(const SegArgs<double, double> &a, bool stage, bool wide_out, hipStream_t st);
hipError_t launch_seg_f32(const SegArgs<float, double> &a, bool stage, bool wide_out, hipStream_t st);

}  // namespace spx
