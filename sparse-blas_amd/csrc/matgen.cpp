// matgen.cpp -- deterministic synthetic CSR matrices for the SpMV workloads.
//
// Own RNG (splitmix64-seeded xoshiro256**) and own samplers, so a (kind,
// sizes, seed) triple gives the same matrix on any machine and any thread
// count: every row draws from its own stream seeded by splitmix64(seed ^ row).
// Two phases: spx_gen_rowlens() fills per-row lengths (the caller turns them
// into a row pointer), spx_gen_fill() fills sorted, distinct columns and
// values row by row.
//
// Also holds the reference harness's 'g' generator
// (SB/test/dspmv_test.cu:137-199), which draws values from glibc rand().
#include "spx_gen.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t t = seed;
    for (auto &w : s) w = splitmix64(t);
  }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform in [0, 1) with 53 random bits
  double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
  // uniform integer in [0, n)
  uint64_t below(uint64_t n) {
    // Lemire's nearly divisionless method (exact, unbiased)
    unsigned __int128 mlt = (unsigned __int128)next() * n;
    uint64_t lo = (uint64_t)mlt;
    if (lo < n) {
      uint64_t thr = (0 - n) % n;
      while (lo < thr) {
        mlt = (unsigned __int128)next() * n;
        lo = (uint64_t)mlt;
      }
    }
    return (uint64_t)(mlt >> 64);
  }
  double normal() {  // Box-Muller, one value per call
    double u1 = u01(), u2 = u01();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

inline uint64_t row_seed(uint64_t seed, uint64_t salt, int64_t row) {
  uint64_t t = seed ^ (salt * 0xD1B54A32D192ED03ull) ^ ((uint64_t)row * 0x9E3779B97F4A7C15ull);
  return splitmix64(t);
}

// Truncated Zipf(s) on [1, kmax] by inverse CDF over a cumulative table.
struct Zipf {
  std::vector<double> cdf;
  Zipf(double s, int kmax) : cdf(kmax) {
    double acc = 0;
    for (int k = 1; k <= kmax; k++) {
      acc += std::pow((double)k, -s);
      cdf[k - 1] = acc;
    }
    for (auto &c : cdf) c /= acc;
  }
  int draw(Rng &r) const {
    double u = r.u01();
    return (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
  }
  double mean() const {
    double prev = 0, m = 0;
    for (size_t i = 0; i < cdf.size(); i++) {
      m += (double)(i + 1) * (cdf[i] - prev);
      prev = cdf[i];
    }
    return m;
  }
};

inline int geometric_trunc(Rng &r, double p, int lo, int hi) {
  // number of trials to first success, truncated by resampling
  for (;;) {
    double u = r.u01();
    int k = 1 + (int)std::floor(std::log1p(-u) / std::log1p(-p));
    if (k >= lo && k <= hi) return k;
  }
}

// k distinct sorted integers drawn uniformly from [lo, hi).
void distinct_sorted(Rng &r, int64_t lo, int64_t hi, int64_t k, int32_t *out) {
  int64_t range = hi - lo;
  if (k >= range) {
    for (int64_t i = 0; i < range; i++) out[i] = (int32_t)(lo + i);
    return;
  }
  if (k * 8 > range) {
    // selection sampling (Knuth algorithm S): one pass over the range
    int64_t need = k, left = range, o = 0;
    for (int64_t i = 0; i < range && need > 0; i++, left--) {
      if ((int64_t)r.below((uint64_t)left) < need) {
        out[o++] = (int32_t)(lo + i);
        need--;
      }
    }
    return;
  }
  std::vector<int32_t> buf;
  buf.reserve(k + k / 4 + 8);
  while ((int64_t)buf.size() < k) {
    int64_t want = k - (int64_t)buf.size();
    for (int64_t i = 0; i < want; i++) buf.push_back((int32_t)(lo + (int64_t)r.below((uint64_t)range)));
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  }
  // drop the surplus at random positions so the sample stays uniform
  while ((int64_t)buf.size() > k) buf.erase(buf.begin() + (int64_t)r.below(buf.size()));
  std::memcpy(out, buf.data(), sizeof(int32_t) * k);
}

double pget(const double *p, int n, int i, double dflt) { return (p && i < n) ? p[i] : dflt; }

}  // namespace

extern "C" {

double spx_gen_zipf_mean(double s, int kmax) { return Zipf(s, kmax).mean(); }

int64_t spx_gen_rowlens(int kind, int m, int n, uint64_t seed, const double *prm, int nprm,
                        int32_t *lens) {
  int64_t total = 0;
  switch (kind) {
    case SPX_GEN_ZIPF: {  // prm: s, kmin, kmax
      double s = pget(prm, nprm, 0, 1.5);
      int kmax = (int)pget(prm, nprm, 2, 512);
      if (kmax > n) kmax = n;
      Zipf z(s, kmax);
#pragma omp parallel for reduction(+ : total) schedule(static)
      for (int i = 0; i < m; i++) {
        Rng r(row_seed(seed, 1, i));
        lens[i] = z.draw(r);
        total += lens[i];
      }
      break;
    }
    case SPX_GEN_CIRCUIT:
    case SPX_GEN_ASIC: {
      // prm: p_geo, kmax_regular, n_dense, dense_lo, dense_hi, dense_total, pinned_len, col_window
      double pg = pget(prm, nprm, 0, 0.1);
      int kmax = (int)pget(prm, nprm, 1, 12);
      int nd = (int)pget(prm, nprm, 2, 200);
      double dlo = pget(prm, nprm, 3, 1e4), dhi = pget(prm, nprm, 4, 1.3e6);
      double dtot = pget(prm, nprm, 5, 0);
      int64_t pinned = (int64_t)pget(prm, nprm, 6, dhi);
      if (nd > m) nd = m;
#pragma omp parallel for schedule(static)
      for (int i = 0; i < m; i++) {
        Rng r(row_seed(seed, 2, i));
        lens[i] = geometric_trunc(r, pg, 1, kmax);
      }
      // dense rows: nd distinct row indices, log-uniform lengths rescaled so they sum to dtot
      Rng rr(row_seed(seed, 3, 0));
      std::vector<int32_t> rows(nd);
      distinct_sorted(rr, 0, m, nd, rows.data());
      // shuffle row order so the pinned row is not always the first
      std::vector<double> dl(nd);
      for (int d = 0; d < nd; d++) dl[d] = std::exp(std::log(dlo) + rr.u01() * (std::log(dhi) - std::log(dlo)));
      if (nd > 0) dl[(size_t)rr.below(nd)] = -1;  // the pinned row
      if (dtot > 0 && nd > 0) {
        double rest = dtot - (double)pinned, sum = 0;
        for (double v : dl) if (v > 0) sum += v;
        double scale = (sum > 0 && rest > 0) ? rest / sum : 1.0;
        for (auto &v : dl) if (v > 0) v = std::min(std::max(v * scale, dlo), (double)n);
      }
      for (int d = 0; d < nd; d++) {
        int64_t L = dl[d] < 0 ? pinned : (int64_t)std::llround(dl[d]);
        if (L > n) L = n;
        lens[rows[d]] = (int32_t)L;
      }
      for (int i = 0; i < m; i++) total += lens[i];
      break;
    }
    case SPX_GEN_LOGNORMAL: {  // prm: mu, sigma, kmin, kmax
      double mu = pget(prm, nprm, 0, 2.0), sg = pget(prm, nprm, 1, 1.0);
      int kmin = (int)pget(prm, nprm, 2, 1);
      int64_t kmax = (int64_t)pget(prm, nprm, 3, n);
      if (kmax > n) kmax = n;
#pragma omp parallel for reduction(+ : total) schedule(static)
      for (int i = 0; i < m; i++) {
        Rng r(row_seed(seed, 4, i));
        double v = std::exp(mu + sg * r.normal());
        int64_t k = (int64_t)std::llround(v);
        if (k < kmin) k = kmin;
        if (k > kmax) k = kmax;
        lens[i] = (int32_t)k;
        total += k;
      }
      break;
    }
    case SPX_GEN_BANDED: {  // prm: bandwidth
      int bw = (int)pget(prm, nprm, 0, 5);
      int h = bw / 2;
#pragma omp parallel for reduction(+ : total) schedule(static)
      for (int i = 0; i < m; i++) {
        int64_t lo = std::max<int64_t>(0, (int64_t)i - h), hi = std::min<int64_t>(n, (int64_t)i + h + 1);
        lens[i] = (int32_t)std::max<int64_t>(0, hi - lo);
        total += lens[i];
      }
      break;
    }
    case SPX_GEN_BLOCKDIAG: {  // prm: block, k_mean, n_dense, dense_len
      int bs = (int)pget(prm, nprm, 0, 64);
      double km = pget(prm, nprm, 1, 8);
      int nd = (int)pget(prm, nprm, 2, 4);
      int64_t dlen = (int64_t)pget(prm, nprm, 3, n / 4);
      if (dlen > n) dlen = n;
#pragma omp parallel for schedule(static)
      for (int i = 0; i < m; i++) {
        Rng r(row_seed(seed, 5, i));
        int64_t b0 = ((int64_t)i / bs) * bs, b1 = std::min<int64_t>(n, b0 + bs);
        int64_t cap = std::max<int64_t>(1, b1 - std::min<int64_t>(b0, n));
        if (b0 >= n) cap = std::min<int64_t>(n, bs);
        int k = geometric_trunc(r, 1.0 / std::max(1.0, km), 1, (int)std::min<int64_t>(cap, 1 << 20));
        lens[i] = k;
      }
      Rng rr(row_seed(seed, 6, 0));
      for (int d = 0; d < nd && m > 0; d++) lens[(size_t)rr.below(m)] = (int32_t)dlen;
      for (int i = 0; i < m; i++) total += lens[i];
      break;
    }
    default:
      return -1;
  }
  return total;
}

int spx_gen_fill(int kind, int m, int n, uint64_t seed, const double *prm, int nprm,
                 const int64_t *rowptr, int32_t *col, double *val) {
  const int vlaw = (int)pget(prm, nprm, 15, -1);  // optional override of the value law
  int window = 0;
  int dense_cut = 1 << 30;
  int vdefault = SPX_VAL_UNIFORM;
  switch (kind) {
    case SPX_GEN_ZIPF: case SPX_GEN_LOGNORMAL: break;
    case SPX_GEN_CIRCUIT:
      window = (int)pget(prm, nprm, 7, 0);
      dense_cut = (int)pget(prm, nprm, 1, 12);
      vdefault = SPX_VAL_NORMAL;
      break;
    case SPX_GEN_ASIC:
      window = (int)pget(prm, nprm, 7, 0);
      dense_cut = (int)pget(prm, nprm, 1, 12);
      vdefault = SPX_VAL_SIGNED_LOGNORMAL;
      break;
    case SPX_GEN_BANDED: case SPX_GEN_BLOCKDIAG: break;
    default: return -1;
  }
  const int law = vlaw >= 0 ? vlaw : vdefault;
  const int bw = (int)pget(prm, nprm, 0, 5), bs = (int)pget(prm, nprm, 0, 64);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < m; i++) {
    Rng r(row_seed(seed, 7, i));
    int64_t b = rowptr[i], k = rowptr[i + 1] - rowptr[i];
    int32_t *c = col + b;
    if (k <= 0) continue;
    if (kind == SPX_GEN_BANDED) {
      int64_t lo = std::max<int64_t>(0, (int64_t)i - bw / 2);
      for (int64_t j = 0; j < k; j++) c[j] = (int32_t)(lo + j);
    } else if (kind == SPX_GEN_BLOCKDIAG && k < n / 2 + 1 && k <= bs) {
      int64_t b0 = std::min<int64_t>(((int64_t)i / bs) * bs, std::max<int64_t>(0, n - bs));
      int64_t b1 = std::min<int64_t>(n, b0 + bs);
      distinct_sorted(r, b0, b1, k, c);
    } else if (window > 0 && k <= dense_cut) {
      // near-diagonal columns: a window of +-window around the row, clipped
      int64_t lo = std::max<int64_t>(0, (int64_t)i - window);
      int64_t hi = std::min<int64_t>(n, (int64_t)i + window + 1);
      if (hi - lo < k) { lo = std::max<int64_t>(0, hi - k); hi = std::min<int64_t>(n, lo + k); }
      distinct_sorted(r, lo, hi, k, c);
    } else {
      distinct_sorted(r, 0, n, k, c);
    }
    double *v = val + b;
    for (int64_t j = 0; j < k; j++) {
      switch (law) {
        case SPX_VAL_UNIFORM: v[j] = 2.0 * r.u01() - 1.0; break;
        case SPX_VAL_NORMAL: v[j] = r.normal(); break;
        case SPX_VAL_ONES: v[j] = 1.0; break;
        case SPX_VAL_SIGNED_LOGNORMAL: {
          double a = std::exp(2.0 * r.normal());
          v[j] = (r.next() & 1) ? -a : a;
          break;
        }
        case SPX_VAL_SMALLINT: v[j] = (double)(int)r.below(10); break;
        default: v[j] = 0; break;
      }
    }
  }
  return 0;
}

void spx_gen_vector(int64_t n, uint64_t seed, int law, double *x) {
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < n; b += 4096) {
    Rng r(row_seed(seed, 8, b));
    int64_t e = std::min<int64_t>(n, b + 4096);
    for (int64_t i = b; i < e; i++) {
      switch (law) {
        case SPX_VAL_NORMAL: x[i] = r.normal(); break;
        case SPX_VAL_ONES: x[i] = 1.0; break;
        case SPX_VAL_SMALLINT: x[i] = (double)(int)r.below(10); break;
        default: x[i] = 2.0 * r.u01() - 1.0; break;
      }
    }
  }
}

int64_t spx_gen_g(int n, unsigned seed, int32_t *coo_row, int32_t *coo_col, double *coo_val) {
  // SB/test/dspmv_test.cu:137-199: rows [0, n/8) hold ceil(0.9 n) entries,
  // the rest ceil(0.01 n); columns 0..k-1; values rand()/RAND_MAX in row order.
  if (n < 8 || n % 8) return -1;
  const int per_block = n / 8;
  int64_t p = 0;
  if (coo_row) srand(seed);
  for (int blk = 0; blk < 8; blk++) {
    const double r = blk == 0 ? 0.9 : 0.01;
    int k = 0;
    while (k < n * r) k++;  // the harness's own bound: j < n*r in double
    for (int ii = blk * per_block; ii < (blk + 1) * per_block; ii++) {
      if (coo_row) {
        for (int j = 0; j < k; j++) {
          coo_row[p + j] = ii;
          coo_col[p + j] = j;
          coo_val[p + j] = (double)rand() / (RAND_MAX);
        }
      }
      p += k;
    }
  }
  return p;
}

}  // extern "C"
