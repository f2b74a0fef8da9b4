/*
 * spx_capi.h -- extern "C" ABI of libspx.so, the MI355X (gfx950) FP64 CSR
 * SpMV library.  Plain pointers and sizes only.  The drop-in C++ functions
 * of spmv_kernel.h are one-line wrappers over the spx_dropin_* calls here;
 * an FFI (ctypes, cgo, JNI) binds these names directly.
 *
 * Two layers:
 *  1. spx_dropin_*: the reference's entry points with identical arguments and
 *     return codes (host pointers in, host y out; stateless per call).
 *  2. spx_plan_*: a device-resident handle (convert once, execute many) that
 *     the drop-in layer is built on.  x and y are DEVICE pointers here.
 *
 * Status codes of the plan layer: SPX_OK (0) or a negative SPX_E* value; the
 * text of the last error on the calling thread is in spx_last_error().
 */
#ifndef SPX_CAPI_H
#define SPX_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SPX_OK = 0,
  SPX_EARG = -1,     /* invalid argument */
  SPX_EMEM = -2,     /* device allocation failed */
  SPX_EDEV = -3,     /* HIP runtime / launch failure */
  SPX_EGUARD = -4,   /* footprint above 0.8 of free device memory */
  SPX_ERANGE = -5    /* size outside what the kernels support (nnz >= 2^31) */
};

/* Work decomposition ("mode") of a plan. */
enum {
  SPX_MODE_ROWS = 0,   /* contiguous row blocks (spMV_mgpu_baseline) */
  SPX_MODE_NNZ = 1,    /* nnz-balanced partitions, one HIP stream each (v1) */
  SPX_MODE_QUEUE = 2   /* nb-nonzero tasks drained from a device work queue
                          by persistent workgroups on q streams (v2) */
};

/* Kernel versions, as in the reference's `kernel` argument. */
enum {
  SPX_KERNEL_VECTOR = 1, /* wavefront/lane-group per row CSR-vector */
  SPX_KERNEL_MERGE = 2,  /* merge-path over (rows + nonzeros) */
  SPX_KERNEL_TILE = 3,   /* CSR5-style 64-lane tiles with bit-flag segments */
  SPX_KERNEL_SEG = 4     /* plan API only: row-aligned units of 8-aligned
                            nonzeros, 16-byte loads, no carries */
};

/* Value storage. */
enum {
  SPX_PREC_F64 = 0,      /* FP64 values, FP64 x */
  SPX_PREC_F32 = 1       /* FP32 values; x stays FP64 in memory and each
                            product splits it into FP32 (hi, lo) in
                            registers; FP64 accumulation, carries and beta
                            update */
};

typedef struct spx_plan_desc {
  int m, n;
  int64_t nnz;
  const int64_t *rowptr64;  /* host row pointer (m+1), or NULL ... */
  const int32_t *rowptr32;  /* ... host int32 row pointer (m+1) */
  const int32_t *col;       /* host column indices (nnz) */
  const double *val;        /* host values (nnz) */
  int kernel;               /* SPX_KERNEL_* */
  int mode;                 /* SPX_MODE_* */
  int parts;                /* ROWS/NNZ: number of partitions (streams), >= 1 */
  int64_t nb;               /* QUEUE: task size in nonzeros (> 0) */
  int queues;               /* QUEUE: streams (copy_of_workspace), >= 1 */
  int sigma;                /* TILE: elements per lane, 4..32 (0 = auto) */
  int lanes;                /* VECTOR: lanes per row, power of 2 <= 64 (0 = auto) */
  int items;                /* MERGE: merge items per thread, 2..16 (0 = auto);
                               SEG: chunks of 2048 nonzeros per unit, 1..32 */
  int precision;            /* SPX_PREC_* */
  int flags;                /* SPX_FLAG_* bits */
  int64_t long_row;         /* rows with at least this many nonzeros take the
                               column-blocked long-row path (x staged in LDS);
                               0 = automatic, < 0 = never */
} spx_plan_desc;

/* desc.flags: the row pointer, column and value arrays are DEVICE memory
 * (the plan copies them; the caller keeps ownership). */
#define SPX_FLAG_DEVICE_PTRS 1
/* desc.flags: kernel-4 FP64 plans run spx_plan_execute_multi as fused passes
 * of up to 64 columns (A read once per pass and held in registers while the
 * columns of X go by; each column of Y is bitwise what a separate execute
 * gives); without it every column is a separate execute. */
#define SPX_FLAG_FUSED_SPMM 2

typedef struct spx_plan *spx_plan_t;

/* Initialise the desc with defaults (kernel 2, one partition, auto tiles). */
void spx_plan_desc_init(spx_plan_desc *d);

/* Upload A to device memory and run the format conversion on the device.
 * Returns SPX_OK and *out, or an error (nothing leaks on failure). */
int spx_plan_create(spx_plan_t *out, const spx_plan_desc *d);
int spx_plan_destroy(spx_plan_t p);

/* y = alpha*A*x + beta*y on device vectors x (n) and y (m), enqueued on
 * `stream` (a hipStream_t; NULL = the null stream).  beta == 0 never reads y.
 * Partitions run on plan-owned streams that fork from and join back to
 * `stream`.  Asynchronous: returns after enqueueing. */
int spx_plan_execute(spx_plan_t p, double alpha, const double *x, double beta,
                     double *y, void *stream);

/* Multi-RHS: Y(:,j) = alpha*A*X(:,j) + beta*Y(:,j) for j < k, X (n-by-k) and
 * Y (m-by-k) column-major with leading dimensions ldx >= n, ldy >= m (see
 * SPX_FLAG_FUSED_SPMM).  A fused pass covers up to 64 columns and reads A
 * once; with long rows it keeps a plan-owned scratch of one double per long-row
 * slice and column (allocated on first use). */
int spx_plan_execute_multi(spx_plan_t p, int k, double alpha, const double *X,
                           int64_t ldx, double beta, double *Y, int64_t ldy,
                           void *stream);

/* Plan facts: writes up to `cap` int64 values:
 * [0] m [1] n [2] nnz [3] kernel [4] mode [5] parts [6] sigma [7] lanes
 * [8] items [9] units (workgroup work items) [10] carries [11] tiles
 * [12] device bytes held [13] precision [14] queues [15] long rows
 * [16] long-row nonzeros [17] column blocks [18] carry runs
 * [19] long-row threshold [20] segmented units that stage their x window
 * [21] 1 when staged and unstaged units run as two concurrent kernels
 * [22] host microseconds spent copying the input's columns and values to
 * the device during the build (kernels 1-3) [23] host microseconds of the
 * whole build [24] column bands of kernel 4's regular rows (0: row units).
 * Returns the count. */
int spx_plan_info(spx_plan_t p, int64_t *out, int cap);

/* Copy plan internals to host for bit-exact checks.  `what`:
 * 0 partitions: 6 int64 per partition (start_idx, end_idx, start_row,
 *   end_row, start_flag, end_flag);
 * 1 tile pointers (uint32, per partition p_i + 1 entries, concatenated);
 * 2 tile flags (uint32, 64 per tile, concatenated);
 * 3 tile-layout column indices (int32); 4 tile-layout values (double);
 * 5 merge-path unit coordinates (int64 pairs: row, nonzero per boundary).
 * Returns the number of BYTES available; copies min(bytes, cap) into dst
 * when dst != NULL. */
int64_t spx_plan_export(spx_plan_t p, int what, void *dst, int64_t cap);

/* Average time (ms) of each of `reps` executions after `warmup` untimed
 * ones, enqueued back-to-back on one stream with HIP events around each.
 * out_ms[0] = whole execute (events around each execute), out_ms[1] = the
 * SpMV kernels alone (the partition setup, fix-up and long-row fold
 * excluded), out_ms[2] = host wall clock from the first timed enqueue to the
 * end of the last execute, per execute.  out_ms holds 3.
 * flush_bytes > 0: before every timed execute a buffer of that size is read,
 * which evicts L2 and MALL and leaves no dirty lines behind; flush_bytes < 0:
 * a buffer of -flush_bytes is overwritten instead (its dirty lines are then
 * written back while the next execute runs).  The flush lies outside the
 * events and inside the wall clock. */
int spx_plan_time(spx_plan_t p, double alpha, const double *x, double beta,
                  double *y, int warmup, int reps, int64_t flush_bytes,
                  double *out_ms);
/* The same timing with every execute's figures kept: exec_ms[i] and
 * kernel_ms[i] (arrays of `reps`) as out_ms[0] and [1] above, *wall_ms (may
 * be NULL) the host wall clock per execute.  With exec_ms and kernel_ms both
 * NULL no events are recorded (each event pair costs the stream a few
 * microseconds) and only *wall_ms is returned: the executes' own rate. */
int spx_plan_time_each(spx_plan_t p, double alpha, const double *x, double beta,
                       double *y, int warmup, int reps, int64_t flush_bytes,
                       double *exec_ms, double *kernel_ms, double *wall_ms);

/* ---- batched form -----------------------------------------------------------
 * Many independent SpMVs y_i = alpha*A_i*x_i + beta*y_i as ONE plan: the
 * matrices are packed block-diagonally (columns shifted by the x offsets),
 * so one launch (persistent workgroups in SPX_MODE_QUEUE) covers them all.
 * x_off/y_off (count+1 entries each, may be NULL) receive where x_i and y_i
 * start inside the concatenated x (sum n_i) and y (sum m_i). */
typedef struct spx_batch_item {
  int m, n;
  int64_t nnz;
  const int64_t *rowptr;    /* host, m+1 */
  const int32_t *col;       /* host, nnz, 0-based within A_i */
  const double *val;        /* host, nnz */
} spx_batch_item;

int spx_batch_create(spx_plan_t *out, int count, const spx_batch_item *items,
                     const spx_plan_desc *tmpl, int64_t *x_off, int64_t *y_off);

/* ---- HIP graphs ---------------------------------------------------------------
 * Capture one execute of a plan (fork/join streams, kernels, fix-up) into an
 * instantiated hipGraph with alpha, beta, x and y baked in; replay it with
 * spx_graph_launch.  Nothing in the graph is a memset node (counters are
 * cleared by a kernel). */
typedef struct spx_graph *spx_graph_t;
int spx_graph_capture(spx_plan_t p, double alpha, const double *x, double beta,
                      double *y, spx_graph_t *out);
int spx_graph_launch(spx_graph_t g, void *stream);
int spx_graph_destroy(spx_graph_t g);
/* Average ms per replay of `reps` back-to-back launches on a private stream
 * (events around the whole run), after `warmup` replays. */
int spx_graph_time(spx_graph_t g, int warmup, int reps, double *out_ms);

/* ---- Matrix Market input (host) ----------------------------------------------------
 * Replaces the reference's loader (main.cu:164-319; SB/include/mmio.h:254-545;
 * SB/test/dspmv_test.cu:101-136).  Reads a coordinate file into a 0-based CSR
 * with an int64 row pointer, keeping the file's order within each row:
 * pattern entries read as 1.0, integer entries as doubles, symmetric and
 * hermitian files mirror each off-diagonal entry, skew-symmetric files mirror
 * it negated.  The three arrays are malloc'd; release each with spx_mm_free.
 * Returns 0, or -1 cannot open, -2 bad banner / unsupported format, -3
 * complex, -4 bad size line, -5 bad entry, -6 out of memory. */
int spx_mm_read(const char *path, int *m, int *n, int64_t *nnz,
                int64_t **rowptr, int32_t **col, double **val);
void spx_mm_free(void *p);

/* ---- learned schedule selector -------------------------------------------------
 * No reference counterpart: it replaces the exhaustive (d, c) search of the
 * reference harness (SB/test/dspmv_test.cu:314-332).  A schedule index s in
 * [0, 96) means kernel 1 + s/24, tile width index (s/6)%4 and partition count
 * {1, 2, 3, 4, 6, 8}[s%6]; widths are lanes {2, 8, 32, 64} (kernel 1), items
 * {2, 4, 8, 16} (kernel 2), sigma {4, 8, 16, 32} (kernel 3), chunks per unit
 * {1, 2, 4, 8} (kernel 4, which ignores the partition count).  All pointers
 * below are device memory. */

/* Parameters of the 32 -> 64 -> 64 -> 96 MLP (FP32, W1 b1 W2 b2 W3 b3 with
 * W[out][in]). */
int spx_sel_param_count(void);

/* The 32 features of a CSR matrix (int32 row pointer and columns on the
 * device), written to feat[32]; rows with at least long_row nonzeros count
 * as long (0 = none).  Synchronous. */
int spx_sel_features(int m, int n, int64_t nnz, const int32_t *rowptr,
                     const int32_t *col, int64_t long_row, float *feat,
                     void *stream);

/* Floats of workspace the MLP entry points need for a batch of B samples. */
int64_t spx_mlp_work_floats(int B);

/* Scores Z[B][96] of feature rows X[B][32]. */
int spx_mlp_forward(const float *params, const float *X, int B, float *Z,
                    float *work, void *stream);

/* One training step on a batch: forward, softmax cross-entropy against the
 * target distributions T[B][96] (loss[B] per sample, may be NULL), gradient
 * of the mean loss into grad[params], then an Adam step (lr, weight decay wd;
 * step counts from 1) on params with moments m1, m2.  lr = 0 skips the
 * update (gradient only). */
int spx_mlp_train_step(float *params, float *m1, float *m2, int step, float lr,
                       float wd, const float *X, const float *T, int B,
                       float *loss, float *grad, float *work, void *stream);

/* out[b] = argmax over allowed schedules (mask[o] != 0; NULL = all) of Z[b]. */
int spx_mlp_argmax(const float *Z, int B, const int32_t *mask, int32_t *out,
                   void *stream);

/* ---- reference drop-in entry points (host pointers) ----------------------
 * Arguments and return codes as SparseBLAS/include/spmv_kernel.h:4-34. */
int spx_dropin_baseline(int m, int n, long long nnz, double *alpha,
                        double *csrVal, long long *csrRowPtr,
                        int *csrColIndex, double *x, double *beta, double *y,
                        int ngpu);
int spx_dropin_v1(int m, int n, long long nnz, double *alpha, double *csrVal,
                  long long *csrRowPtr, int *csrColIndex, double *x,
                  double *beta, double *y, int ngpu, int kernel);
int spx_dropin_v2(int m, int n, long long nnz, double *alpha, double *csrVal,
                  long long *csrRowPtr, int *csrColIndex, double *x,
                  double *beta, double *y, int ngpu, int kernel, long long nb,
                  int copy_of_workspace);
/* Device pointers, int32 row pointer (SparseBLAS/src/csr5_kernel.cu:12-32). */
int spx_dropin_csr5_kernel(int m, int n, int nnz, double *alpha,
                           double *csrVal, int *csrRowPtr, int *csrColIndex,
                           double *x, double *beta, double *y);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The arguments and result of spx_dropin_csr5_kernel
 * (device pointers, int32 CSR, y = alpha*A*x + beta*y), computed by the
 * merge-path kernel directly on the caller's arrays after a device check of
 * the row pointer and column indices.  Its unit table and carry slots live
 * in a file-scope device array of 8 MiB per GPU (address taken once per
 * device), so m + nnz may be at most 306,774,016; larger inputs fail with -1
 * and a message in spx_last_error.  Calls that overlap on one GPU share the
 * array and race: serialise them.  Synchronous; returns 0 or -1. */
int spx_dropin_csr5_kernel_noalloc(int m, int n, int nnz, double *alpha,
                                   double *csrVal, int *csrRowPtr,
                                   int *csrColIndex, double *x, double *beta,
                                   double *y);
int spx_get_row_from_index(int n, long long *a, long long idx);
double spx_get_time(void);
double spx_get_gpu_availble_mem(int ngpu);

/* Phases of the calling thread's last spx_dropin_baseline / v1 / v2 call, in
 * milliseconds of host time: [0] plan build (including the upload of the
 * matrix) [1] of which the matrix upload [2] upload of x (and y when
 * beta != 0) [3] SpMV execute [4] download of y [5] release of device
 * memory [6] whole call.  Returns the count written (at most cap). */
int spx_dropin_last_times(double *ms, int cap);

/* Number of stream partitions the v1 drop-in uses at ngpu = 1 (default 1;
 * the schedule selector sets it).  Returns the previous value. */
int spx_set_v1_streams(int parts);

/* Host-side seam for tests: when bytes >= 0, the memory guard uses it as the
 * free device memory instead of hipMemGetInfo (-1 restores the query). */
void spx_debug_set_free_mem(long long bytes);

/* Host-side seam for tests: when after >= 0, that many more device
 * allocations of the library succeed and the next one fails as out of
 * memory (the drop-ins then return 1, as the reference does on a failed
 * cudaMalloc, SparseBLAS/src/dspmv_mgpu_baseline.cu:119-153); -1 disarms. */
void spx_debug_fail_alloc(long long after);

/* ---- device memory helpers (so FFI callers need no HIP bindings) -------- */
void *spx_malloc(int64_t bytes);
int spx_free(void *p);
int spx_memcpy_h2d(void *dst, const void *src, int64_t bytes);
int spx_memcpy_d2h(void *dst, const void *src, int64_t bytes);
int spx_memcpy_d2d(void *dst, const void *src, int64_t bytes);
int spx_synchronize(void);
int spx_device_count(void);

const char *spx_last_error(void);
const char *spx_version(void);

#ifdef __cplusplus
}
#endif
#endif
