// mmio.cpp -- Matrix Market coordinate files to CSR (host only).
//
// Follows the reference's loader (main.cu:164-319, with the banner and size
// parsing of SB/include/mmio.h:254-545): 1-based indices become 0-based,
// `pattern` entries read as 1.0 (the reference harness's binary mode wrote
// 1e-5, SB/test/dspmv_test.cu:123-125; 1.0 keeps pattern matrices exact in
// FP32 storage), `integer` entries convert to double, `symmetric` and
// `hermitian` files mirror every off-diagonal entry, and the CSR keeps the
// file's order within a row (a counting sort by row, main.cu:262-312).
// `skew-symmetric` files mirror with the sign flipped (the reference reads
// them as general, dropping the mirrored half).  Complex files and the dense
// `array` format are rejected.  Return codes follow main.cu: -1 cannot open,
// -2 bad banner, -3 complex, -4 bad size line; -5 bad entry (index out of
// range or a short line), -6 out of memory.
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spx_capi.h"

namespace {

std::string lower(std::string s) {
  for (char &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

// Reads the next line that is not a comment or blank; false at end of file.
bool next_line(FILE *f, std::string &line) {
  char buf[4096];
  while (std::fgets(buf, sizeof buf, f)) {
    line = buf;
    while (!line.empty() && line.back() != '\n' && !std::feof(f)) {  // long line: keep reading
      if (!std::fgets(buf, sizeof buf, f)) break;
      line += buf;
    }
    size_t i = line.find_first_not_of(" \t\r\n");
    if (i == std::string::npos || line[i] == '%') continue;
    return true;
  }
  return false;
}

}  // namespace

extern "C" {

int spx_mm_read(const char *path, int *m_out, int *n_out, int64_t *nnz_out, int64_t **rowptr_out,
                int32_t **col_out, double **val_out) {
  if (!path || !m_out || !n_out || !nnz_out || !rowptr_out || !col_out || !val_out) return -1;
  *rowptr_out = nullptr;
  *col_out = nullptr;
  *val_out = nullptr;
  FILE *f = std::fopen(path, "r");
  if (!f) return -1;
  char banner[1024];
  if (!std::fgets(banner, sizeof banner, f)) {
    std::fclose(f);
    return -2;
  }
  char t0[64] = {0}, t1[64] = {0}, t2[64] = {0}, t3[64] = {0}, t4[64] = {0};
  if (std::sscanf(banner, "%63s %63s %63s %63s %63s", t0, t1, t2, t3, t4) != 5 ||
      std::strcmp(t0, "%%MatrixMarket") != 0 || lower(t1) != "matrix") {
    std::fclose(f);
    return -2;
  }
  const std::string fmt = lower(t2), field = lower(t3), sym = lower(t4);
  if (field == "complex") {
    std::fclose(f);
    return -3;
  }
  const bool pattern = field == "pattern", integer = field == "integer", real = field == "real" || field == "double";
  const bool symmetric = sym == "symmetric" || sym == "hermitian", skew = sym == "skew-symmetric";
  if (fmt != "coordinate" || !(pattern || integer || real) || !(symmetric || skew || sym == "general")) {
    std::fclose(f);
    return -2;
  }
  std::string line;
  long long M = 0, N = 0, L = 0;
  if (!next_line(f, line) || std::sscanf(line.c_str(), "%lld %lld %lld", &M, &N, &L) != 3 || M < 0 || N < 0 ||
      L < 0 || M > INT32_MAX || N > INT32_MAX) {
    std::fclose(f);
    return -4;
  }
  std::vector<int32_t> ri, ci;
  std::vector<double> vv;
  try {
    ri.resize((size_t)L);
    ci.resize((size_t)L);
    vv.resize((size_t)L);
  } catch (...) {
    std::fclose(f);
    return -6;
  }
  for (long long k = 0; k < L; k++) {
    long long i = 0, j = 0;
    double v = 1.0;
    int got = 0;
    if (next_line(f, line)) {
      if (pattern) {
        got = std::sscanf(line.c_str(), "%lld %lld", &i, &j) == 2;
      } else if (integer) {
        long long iv = 0;
        got = std::sscanf(line.c_str(), "%lld %lld %lld", &i, &j, &iv) == 3;
        v = (double)iv;
      } else {
        got = std::sscanf(line.c_str(), "%lld %lld %lg", &i, &j, &v) == 3;
      }
    }
    if (!got || i < 1 || j < 1 || i > M || j > N || ((symmetric || skew) && (i > N || j > M))) {
      std::fclose(f);
      return -5;
    }
    ri[(size_t)k] = (int32_t)(i - 1);
    ci[(size_t)k] = (int32_t)(j - 1);
    vv[(size_t)k] = v;
  }
  std::fclose(f);

  // counting sort by row; a mirrored entry goes to its row right after the
  // original, both in file order
  const bool mirror = symmetric || skew;
  std::vector<int64_t> cnt((size_t)M + 1, 0);
  for (long long k = 0; k < L; k++) {
    cnt[(size_t)ri[(size_t)k] + 1]++;
    if (mirror && ri[(size_t)k] != ci[(size_t)k]) cnt[(size_t)ci[(size_t)k] + 1]++;
  }
  for (long long r = 0; r < M; r++) cnt[(size_t)r + 1] += cnt[(size_t)r];
  const int64_t nnz = cnt[(size_t)M];
  int64_t *rp = (int64_t *)std::malloc(sizeof(int64_t) * ((size_t)M + 1));
  int32_t *col = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  double *val = (double *)std::malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
  if (!rp || !col || !val) {
    std::free(rp);
    std::free(col);
    std::free(val);
    return -6;
  }
  std::memcpy(rp, cnt.data(), sizeof(int64_t) * ((size_t)M + 1));
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  for (long long k = 0; k < L; k++) {
    const int32_t i = ri[(size_t)k], j = ci[(size_t)k];
    const double v = vv[(size_t)k];
    int64_t at = fill[(size_t)i]++;
    col[at] = j;
    val[at] = v;
    if (mirror && i != j) {
      at = fill[(size_t)j]++;
      col[at] = i;
      val[at] = skew ? -v : v;
    }
  }
  *m_out = (int)M;
  *n_out = (int)N;
  *nnz_out = nnz;
  *rowptr_out = rp;
  *col_out = col;
  *val_out = val;
  return 0;
}

void spx_mm_free(void *p) { std::free(p); }

}  // extern "C"
