/*
 * oracle.h -- CPU restatement of Sparse-BLAS's host-side algorithms for the
 * FP64 CSR SpMV path.  TEST INFRASTRUCTURE ONLY: only tests/, the smoke check
 * in __graft_entry__.py and the cpu_baseline leg of bench.py may load this
 * library, and only as the checker.  The product never links it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference root; SB/ = SparseBLAS/).
 *
 * Parity status of each function (how it is pinned) is listed in
 * oracle/README.md.
 */
#ifndef SPX_ORACLE_H
#define SPX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- SpMV --------------------------------------------------------------- */

/* main.cu:350-356: per row, sum = 0; sum += x[col]*val*alpha (left to right);
 * then the BLAS beta term y = sum + beta*y0 (beta == 0 does not read y), the
 * contract cusparseDcsrmv is called with (SB/src/dspmv_mgpu_v1.cu:200-204). */
void orc_spmv(int m, const int64_t *rowptr, const int32_t *col,
              const double *val, const double *x, double alpha, double beta,
              double *y);
/* The same loop, rows split statically over nthreads OpenMP threads. */
void orc_spmv_omp(int m, const int64_t *rowptr, const int32_t *col,
                  const double *val, const double *x, double alpha,
                  double beta, double *y, int nthreads);

/* Same loop over FP32-stored values (the precision form's reference is still
 * FP64: each value is widened to double before the product). */
void orc_spmv_f32val(int m, const int64_t *rowptr, const int32_t *col,
                     const float *val, const double *x, double alpha,
                     double beta, double *y);

/* Compensated (Ogita-Rump-Oishi Dot2) row sums: y = alpha*(sum v*x) + beta*y0
 * evaluated in double-double, then rounded once.  Used to bound the error of
 * the plain oracle itself; also writes the per-row magnitude |A||x| into
 * absax (may be NULL). */
void orc_spmv_dot2(int m, const int64_t *rowptr, const int32_t *col,
                   const double *val, const double *x, double alpha,
                   double beta, double *y, double *absax);

/* Multi-RHS: Y(:,j) = alpha*A*X(:,j) + beta*Y(:,j), X n-by-k and Y m-by-k,
 * both column-major.  Column j is exactly orc_spmv on X(:,j). */
void orc_spmm(int m, int n, int k, const int64_t *rowptr, const int32_t *col,
              const double *val, const double *X, double alpha, double beta,
              double *Y);

/* Same loop, rows [r0, r1) only (bounded CPU baseline samples). */
void orc_spmv_rows(int r0, int r1, const int64_t *rowptr, const int32_t *col,
                   const double *val, const double *x, double alpha,
                   double beta, double *y);

/* ---- row search / partitions ------------------------------------------- */

/* SB/src/spmv_helper.cu:16-39, restated with its behaviour on runs of equal
 * row pointers (empty rows): it returns the probe index that matches first. */
int orc_get_row_from_index(int n, const int64_t *a, int64_t idx);

/* The row holding nonzero k: upper_bound(rowptr[0..m], k) - 1.  This is the
 * rule the CSR5 tile pointer uses (SB/include/detail/cuda/format_cuda.h:41). */
int orc_row_of(int m, const int64_t *rowptr, int64_t k);

/* SB/src/dspmv_mgpu_v1.cu:59-94: nnz-balanced split into P parts.
 * start = floor((double)(i*nnz)/P), end = floor((double)((i+1)*nnz)/P) - 1;
 * rows by `rule` (0 = reference get_row_from_index, 1 = upper_bound-1);
 * start_flag = start > rowptr[start_row]; end_flag = end < rowptr[end_row+1]-1. */
void orc_v1_partition(int m, int64_t nnz, const int64_t *rowptr, int P,
                      int rule, int64_t *start_idx, int64_t *end_idx,
                      int32_t *start_row, int32_t *end_row,
                      uint8_t *start_flag, uint8_t *end_flag);

/* SB/src/dspmv_mgpu_v2.cu:218-267: T = ceil(nnz/nb) tasks with
 * start = t*nnz/T, end = (t+1)*nnz/T - 1 (integer division), rows and flags as
 * in v1.  Returns T; arrays must hold at least T entries (call with NULL
 * arrays to get T). */
int orc_v2_tasks(int m, int64_t nnz, const int64_t *rowptr, int64_t nb,
                 int rule, int64_t *start_idx, int64_t *end_idx,
                 int32_t *start_row, int32_t *end_row, uint8_t *start_flag,
                 uint8_t *end_flag);

/* The whole v1 drop-in at ngpu = P on the CPU: per partition, local int32
 * row pointer (v1:125-133), local SpMV y_p = alpha*A_p*x + beta*y_p, then the
 * host fix-up of split rows in partition order (v1:232-245):
 * y[start_row] = y_p[0] + prev - beta*y0[start_row].
 * Uses the upper_bound row rule.  Returns 0. */
int orc_v1_spmv(int m, int64_t nnz, const int64_t *rowptr, const int32_t *col,
                const double *val, const double *x, double alpha, double beta,
                double *y, int P);

/* ---- CSR5-style tiles (64 lanes, sigma elements per lane) ---------------- */

/* SB/include/anonymouslib_cuda.h:297-313 restated: q = nnz/m;
 * sigma = 4 if q <= 4, q if q <= 32, 32 if q <= 256, else 6. */
int orc_tile_sigma(int m, int64_t nnz);

/* Number of tiles p = ceil(nnz / (64*sigma)) (anonymouslib_cuda.h:137 with
 * omega = 64). */
int64_t orc_tile_count(int64_t nnz, int sigma);

/* format_cuda.h:21-95 with omega = 64: tile_ptr[t] = upper_bound(rowptr,
 * min(t*64*sigma, nnz)) - 1 for t in [0, p]; bit 31 of tile_ptr[t] (t < p) is
 * set when some row r in [tile_ptr[t], tile_ptr[t+1]) is empty. */
void orc_tile_ptr(int m, int64_t nnz, const int64_t *rowptr, int sigma,
                  int64_t p, uint32_t *tile_ptr);

/* Row-start bit flags, one 32-bit word per (tile, lane) (sigma <= 32):
 * bit k of flags[t*64 + l] is set when element t*64*sigma + l*sigma + k is the
 * first nonzero of some row (format_cuda.h:140-157, same position rule).
 * Covers every tile, the partial last tile included. */
void orc_tile_flags(int m, int64_t nnz, const int64_t *rowptr, int sigma,
                    int64_t p, uint32_t *flags);

/* AoSoA layout of a tile (format_cuda.h:525-585, R2C direction): element k of
 * lane l of tile t moves from CSR position t*64*sigma + l*sigma + k to
 * t*64*sigma + k*64 + l.  Every tile is transposed, the partial last one
 * included: its positions past nnz are left out of the CSR side and filled
 * with zero on the tile side (the output arrays hold p*64*sigma entries). */
void orc_tile_transpose_i32(int64_t nnz, int sigma, int64_t p,
                            const int32_t *in, int32_t *out);
void orc_tile_transpose_f64(int64_t nnz, int sigma, int64_t p,
                            const double *in, double *out);

/* ---- inputs -------------------------------------------------------------- */

/* SB/test/dspmv_test.cu:137-208, 'g' mode: m = n, nb = m/8; rows [0, nb)
 * hold ceil(0.9 n) nonzeros, the others ceil(0.01 n) (loop `j < n*r`);
 * columns 0..k-1; value = rand()/RAND_MAX from glibc rand() after srand(seed)
 * (the harness never calls srand, so glibc's default seed 1 applies).
 * Requires n % 8 == 0 and n >= 8 (other n overrun the reference's arrays).
 * Returns nnz; with NULL outputs only counts. */
int64_t orc_gen_g(int n, unsigned seed, int32_t *coo_row, int32_t *coo_col,
                  double *coo_val);

/* dspmv_test.cu:228-251: COO (rows in any order) to CSR row pointer by
 * counting; the column/value order is the COO order within each row, which
 * the 'g' generator already emits row-major. */
void orc_coo_to_csr_rowptr(int m, int64_t nnz, const int32_t *coo_row,
                           int64_t *rowptr);

/* dspmv_test.cu:281-282: alpha, beta = the next two rand()/RAND_MAX draws. */
void orc_harness_alpha_beta(unsigned seed, int64_t draws_before,
                            double *alpha, double *beta);

/* SB/include/detail/utils.h:10-20: getB and getFLOP. */
double orc_getB(int m, int64_t nnz);
double orc_getFLOP(int64_t nnz);

#ifdef __cplusplus
}
#endif
#endif
