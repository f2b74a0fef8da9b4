// spx_gen.h -- deterministic synthetic CSR inputs (host only, C ABI).
#ifndef SPX_GEN_H
#define SPX_GEN_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

// Row-length laws.
enum {
  SPX_GEN_ZIPF = 1,       // prm: s, kmin, kmax; uniform distinct columns
  SPX_GEN_CIRCUIT = 2,    // prm: p_geo, kmax, n_dense, dense_lo, dense_hi, dense_total, pinned, col_window
  SPX_GEN_ASIC = 3,       // same parameters as CIRCUIT, signed-lognormal values
  SPX_GEN_LOGNORMAL = 4,  // prm: mu, sigma, kmin, kmax; uniform distinct columns
  SPX_GEN_BANDED = 5,     // prm: bandwidth; consecutive columns around the diagonal
  SPX_GEN_BLOCKDIAG = 6   // prm: block, k_mean, n_dense, dense_len
};
// Value laws (prm[15] overrides the kind's default when >= 0).
enum {
  SPX_VAL_UNIFORM = 0,          // U[-1, 1)
  SPX_VAL_NORMAL = 1,           // N(0, 1)
  SPX_VAL_ONES = 2,             // 1.0
  SPX_VAL_SIGNED_LOGNORMAL = 3, // +-exp(2 N(0,1))
  SPX_VAL_SMALLINT = 4          // integers 0..9
};

double spx_gen_zipf_mean(double s, int kmax);
int64_t spx_gen_rowlens(int kind, int m, int n, uint64_t seed, const double *prm, int nprm,
                        int32_t *lens);
int spx_gen_fill(int kind, int m, int n, uint64_t seed, const double *prm, int nprm,
                 const int64_t *rowptr, int32_t *col, double *val);
void spx_gen_vector(int64_t n, uint64_t seed, int law, double *x);
int64_t spx_gen_g(int n, unsigned seed, int32_t *coo_row, int32_t *coo_col, double *coo_val);

#ifdef __cplusplus
}
#endif
#endif
