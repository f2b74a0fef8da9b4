/*
 * oracle.c -- CPU restatement of Sparse-BLAS's host algorithms for the FP64
 * CSR SpMV path.  TEST INFRASTRUCTURE ONLY (see oracle.h): the checker, never
 * the thing measured or shipped.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- SpMV --------------------------------------------------------------- */

void orc_spmv_rows(int r0, int r1, const int64_t *rowptr, const int32_t *col,
                   const double *val, const double *x, double alpha,
                   double beta, double *y)
{
    /* main.cu:350-356 */
    for (int i = r0; i < r1; i++) {
        double sum = 0;
        for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++)
            sum += x[col[j]] * val[j] * alpha;
        y[i] = (beta == 0.0) ? sum : sum + beta * y[i];
    }
}

void orc_spmv(int m, const int64_t *rowptr, const int32_t *col,
              const double *val, const double *x, double alpha, double beta,
              double *y)
{
    orc_spmv_rows(0, m, rowptr, col, val, x, alpha, beta, y);
}

void orc_spmv_omp(int m, const int64_t *rowptr, const int32_t *col,
                  const double *val, const double *x, double alpha,
                  double beta, double *y, int nthreads)
{
    /* main.cu:350-356 with the rows split statically over nthreads OpenMP
     * threads: every row is summed exactly as orc_spmv sums it, so the
     * result is bitwise the same for any thread count. */
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int i = 0; i < m; i++) {
        double sum = 0;
        for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++)
            sum += x[col[j]] * val[j] * alpha;
        y[i] = (beta == 0.0) ? sum : sum + beta * y[i];
    }
}

void orc_spmv_f32val(int m, const int64_t *rowptr, const int32_t *col,
                     const float *val, const double *x, double alpha,
                     double beta, double *y)
{
    for (int i = 0; i < m; i++) {
        double sum = 0;
        for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++)
            sum += x[col[j]] * (double)val[j] * alpha;
        y[i] = (beta == 0.0) ? sum : sum + beta * y[i];
    }
}

void orc_spmm(int m, int n, int k, const int64_t *rowptr, const int32_t *col,
              const double *val, const double *X, double alpha, double beta,
              double *Y)
{
    for (int j = 0; j < k; j++)
        orc_spmv(m, rowptr, col, val, X + (size_t)j * n, alpha, beta,
                 Y + (size_t)j * m);
}

/* Error-free transformations (Knuth TwoSum, FMA TwoProduct). */
static void two_sum(double a, double b, double *s, double *e)
{
    double t = a + b;
    double z = t - a;
    *e = (a - (t - z)) + (b - z);
    *s = t;
}

void orc_spmv_dot2(int m, const int64_t *rowptr, const int32_t *col,
                   const double *val, const double *x, double alpha,
                   double beta, double *y, double *absax)
{
    for (int i = 0; i < m; i++) {
        double s = 0, c = 0, mag = 0;
        for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++) {
            double p = val[j] * x[col[j]];
            double pe = fma(val[j], x[col[j]], -p);
            double t, e;
            two_sum(s, p, &t, &e);
            s = t;
            c += e + pe;
            mag += fabs(p);
        }
        /* alpha*(s + c) + beta*y0 carried as a double-double */
        double hi = s + c;
        double lo = c - (hi - s);
        double ah = alpha * hi;
        double ae = fma(alpha, hi, -ah) + alpha * lo;
        if (beta != 0.0) {
            double by = beta * y[i];
            double bye = fma(beta, y[i], -by);
            double t, e;
            two_sum(ah, by, &t, &e);
            y[i] = t + (e + ae + bye);
        } else {
            y[i] = ah + ae;
        }
        if (absax)
            absax[i] = fabs(alpha) * mag;
    }
}

/* ---- row search / partitions ------------------------------------------- */

int orc_get_row_from_index(int n, const int64_t *a, int64_t idx)
{
    /* SB/src/spmv_helper.cu:16-39 */
    int l = 0, r = n;
    while (l < r - 1) {
        int mid = l + (r - l) / 2;
        if (idx < a[mid])
            r = mid;
        else if (idx > a[mid])
            l = mid;
        else
            return mid;
    }
    if (idx == a[l])
        return l;
    if (idx == a[r])
        return r;
    return l;
}

int orc_row_of(int m, const int64_t *rowptr, int64_t k)
{
    /* upper_bound over rowptr[0..m] (format_cuda.h:41 with utils_cuda.h:25-53) */
    int lo = 0, hi = m + 1;
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (rowptr[mid] <= k)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo - 1;
}

static int pick_row(int m, const int64_t *rowptr, int64_t k, int rule)
{
    return rule == 0 ? orc_get_row_from_index(m, rowptr, k)
                     : orc_row_of(m, rowptr, k);
}

static void rows_and_flags(int m, const int64_t *rowptr, int rule, int64_t s,
                           int64_t e, int32_t *sr, int32_t *er, uint8_t *sf,
                           uint8_t *ef)
{
    /* SB/src/dspmv_mgpu_v1.cu:72-94 */
    int r0 = pick_row(m, rowptr, s, rule);
    int r1 = pick_row(m, rowptr, e, rule);
    *sr = r0;
    *er = r1;
    *sf = (uint8_t)(s > rowptr[r0]);
    *ef = (uint8_t)(e < rowptr[r1 + 1] - 1);
}

void orc_v1_partition(int m, int64_t nnz, const int64_t *rowptr, int P,
                      int rule, int64_t *start_idx, int64_t *end_idx,
                      int32_t *start_row, int32_t *end_row,
                      uint8_t *start_flag, uint8_t *end_flag)
{
    for (int i = 0; i < P; i++) {
        /* SB/src/dspmv_mgpu_v1.cu:62-69 (double division, then floor) */
        long long tmp1 = (long long)i * nnz;
        long long tmp2 = (long long)(i + 1) * nnz;
        start_idx[i] = (int64_t)floor((double)tmp1 / P);
        end_idx[i] = (int64_t)floor((double)tmp2 / P) - 1;
        rows_and_flags(m, rowptr, rule, start_idx[i], end_idx[i],
                       &start_row[i], &end_row[i], &start_flag[i],
                       &end_flag[i]);
    }
}

int orc_v2_tasks(int m, int64_t nnz, const int64_t *rowptr, int64_t nb,
                 int rule, int64_t *start_idx, int64_t *end_idx,
                 int32_t *start_row, int32_t *end_row, uint8_t *start_flag,
                 uint8_t *end_flag)
{
    /* SB/src/dspmv_mgpu_v2.cu:218 */
    int T = (int)((nnz + nb - 1) / nb);
    if (!start_idx)
        return T;
    for (int t = 0; t < T; t++) {
        /* v2:229-236: integer division inside, then cast and floor */
        long long tmp1 = (long long)t * nnz;
        long long tmp2 = (long long)(t + 1) * nnz;
        start_idx[t] = (int64_t)floor((double)(tmp1 / T));
        end_idx[t] = (int64_t)floor((double)(tmp2 / T)) - 1;
        rows_and_flags(m, rowptr, rule, start_idx[t], end_idx[t],
                       &start_row[t], &end_row[t], &start_flag[t],
                       &end_flag[t]);
    }
    return T;
}

int orc_v1_spmv(int m, int64_t nnz, const int64_t *rowptr, const int32_t *col,
                const double *val, const double *x, double alpha, double beta,
                double *y, int P)
{
    int64_t *s = malloc(sizeof(int64_t) * P), *e = malloc(sizeof(int64_t) * P);
    int32_t *sr = malloc(sizeof(int32_t) * P), *er = malloc(sizeof(int32_t) * P);
    uint8_t *sf = malloc(P), *ef = malloc(P);
    double *y0 = malloc(sizeof(double) * (m > 0 ? m : 1));
    memcpy(y0, y, sizeof(double) * m);
    orc_v1_partition(m, nnz, rowptr, P, 1, s, e, sr, er, sf, ef);
    for (int d = 0; d < P; d++) {
        int dm = er[d] - sr[d] + 1;
        int64_t dnnz = e[d] - s[d] + 1;
        if (dnnz <= 0)
            continue;
        /* v1:125-133 local row pointer */
        int64_t *lrp = malloc(sizeof(int64_t) * (dm + 1));
        lrp[0] = 0;
        lrp[dm] = dnnz;
        for (int j = 1; j < dm; j++)
            lrp[j] = rowptr[sr[d] + j] - s[d];
        double *yl = malloc(sizeof(double) * dm);
        memcpy(yl, y0 + sr[d], sizeof(double) * dm);
        orc_spmv(dm, lrp, col + s[d], val + s[d], x, alpha, beta, yl);
        /* v1:232-245 */
        double prev = 0.0;
        if (sf[d])
            prev = y[sr[d]];
        memcpy(y + sr[d], yl, sizeof(double) * dm);
        if (sf[d]) {
            y[sr[d]] += prev;
            y[sr[d]] -= y0[sr[d]] * beta;
        }
        free(lrp);
        free(yl);
    }
    free(s); free(e); free(sr); free(er); free(sf); free(ef); free(y0);
    return 0;
}

/* ---- CSR5-style tiles -------------------------------------------------- */

#define ORC_OMEGA 64

int orc_tile_sigma(int m, int64_t nnz)
{
    /* SB/include/anonymouslib_cuda.h:297-313 */
    const int r = 4, s = 32, t = 256, u = 6;
    int64_t q = m > 0 ? nnz / m : 0;
    if (q <= r)
        return r;
    if (q <= s)
        return (int)q;
    if (q <= t)
        return s;
    return u;
}

int64_t orc_tile_count(int64_t nnz, int sigma)
{
    int64_t w = (int64_t)ORC_OMEGA * sigma;
    return (nnz + w - 1) / w;
}

void orc_tile_ptr(int m, int64_t nnz, const int64_t *rowptr, int sigma,
                  int64_t p, uint32_t *tile_ptr)
{
    /* format_cuda.h:21-42 */
    for (int64_t t = 0; t <= p; t++) {
        int64_t b = t * ORC_OMEGA * sigma;
        if (b > nnz)
            b = nnz;
        tile_ptr[t] = (uint32_t)orc_row_of(m, rowptr, b);
    }
    /* format_cuda.h:44-95: flag tiles whose rows [start, stop) hold an empty row */
    for (int64_t t = 0; t < p; t++) {
        uint32_t start = tile_ptr[t] & 0x7FFFFFFFu;
        uint32_t stop = tile_ptr[t + 1] & 0x7FFFFFFFu;
        if (start == stop)
            continue;
        for (uint32_t r = start; r < stop; r++) {
            if (rowptr[r] == rowptr[r + 1]) {
                tile_ptr[t] = start | 0x80000000u;
                break;
            }
        }
    }
}

void orc_tile_flags(int m, int64_t nnz, const int64_t *rowptr, int sigma,
                    int64_t p, uint32_t *flags)
{
    memset(flags, 0, sizeof(uint32_t) * (size_t)p * ORC_OMEGA);
    /* format_cuda.h:140-157: every row start sets a bit at its position */
    for (int r = 0; r < m; r++) {
        int64_t off = rowptr[r];
        if (off >= nnz || rowptr[r + 1] == off)
            continue; /* empty rows own no element */
        int64_t gx = off / sigma;
        int64_t lane = gx % ORC_OMEGA;
        int64_t tile = gx / ORC_OMEGA;
        int bit = (int)(off % sigma);
        flags[tile * ORC_OMEGA + lane] |= 1u << bit;
    }
}

#define TRANSPOSE_BODY                                                        \
    int64_t w = (int64_t)ORC_OMEGA * sigma;                                   \
    for (int64_t t = 0; t < p; t++)                                           \
        for (int l = 0; l < ORC_OMEGA; l++)                                   \
            for (int k = 0; k < sigma; k++) {                                 \
                int64_t src = t * w + (int64_t)l * sigma + k;                 \
                out[t * w + (int64_t)k * ORC_OMEGA + l] =                     \
                    src < nnz ? in[src] : 0;                                  \
            }

void orc_tile_transpose_i32(int64_t nnz, int sigma, int64_t p,
                            const int32_t *in, int32_t *out)
{
    TRANSPOSE_BODY
}

void orc_tile_transpose_f64(int64_t nnz, int sigma, int64_t p,
                            const double *in, double *out)
{
    TRANSPOSE_BODY
}

/* ---- inputs -------------------------------------------------------------- */

int64_t orc_gen_g(int n, unsigned seed, int32_t *coo_row, int32_t *coo_col,
                  double *coo_val)
{
    /* SB/test/dspmv_test.cu:137-199 */
    int m = n;
    int nb = m / 8;
    if (nb <= 0 || m % 8 != 0)
        return -1;
    const double r1 = 0.9, r2 = 0.01;
    int64_t p = 0;
    if (coo_row)
        srand(seed);
    for (int i = 0; i < m; i += nb) {
        double r = (i == 0) ? r1 : r2;
        for (int ii = i; ii < i + nb; ii++) {
            for (int j = 0; j < n * r; j++) {
                if (coo_row) {
                    coo_row[p] = ii;
                    coo_col[p] = j;
                    coo_val[p] = (double)rand() / (RAND_MAX);
                }
                p++;
            }
        }
    }
    return p;
}

void orc_coo_to_csr_rowptr(int m, int64_t nnz, const int32_t *coo_row,
                           int64_t *rowptr)
{
    /* SB/test/dspmv_test.cu:228-251 */
    int64_t *counter = calloc((size_t)(m > 0 ? m : 1), sizeof(int64_t));
    for (int64_t i = 0; i < nnz; i++)
        counter[coo_row[i]]++;
    rowptr[0] = 0;
    for (int i = 1; i <= m; i++)
        rowptr[i] = rowptr[i - 1] + counter[i - 1];
    free(counter);
}

void orc_harness_alpha_beta(unsigned seed, int64_t draws_before,
                            double *alpha, double *beta)
{
    srand(seed);
    for (int64_t i = 0; i < draws_before; i++)
        (void)rand();
    *alpha = (double)rand() / (RAND_MAX);
    *beta = (double)rand() / (RAND_MAX);
}

double orc_getB(int m, int64_t nnz)
{
    /* SB/include/detail/utils.h:10-14 with int indices, double values */
    return (double)((int64_t)(m + 1 + nnz) * sizeof(int) +
                    (2 * nnz + m) * sizeof(double));
}

double orc_getFLOP(int64_t nnz)
{
    /* SB/include/detail/utils.h:16-20 */
    return (double)(2 * nnz);
}
