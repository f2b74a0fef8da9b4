// handle_check.cpp -- drives anonymouslibHandle (include/anonymouslib_hip.h)
// the way the reference's csr5_kernel wrapper drives its handle
// (SparseBLAS/src/csr5_kernel.cu:17-30), for tests/test_handle.py.
//
//   handle_check host          host-only logic: auto sigma, error codes
//   handle_check spmv <dir>    inputs from <dir>/*.bin, outputs y*.bin there
//
// Every result is one JSON line on stdout.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../include/anonymouslib_hip.h"

namespace {

template <typename T>
bool read_bin(const std::string &path, std::vector<T> &v, size_t count) {
  std::ifstream f(path, std::ios::binary);
  v.resize(count);
  if (count == 0) return bool(f);
  return bool(f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(count * sizeof(T))));
}

template <typename T>
bool write_bin(const std::string &path, const std::vector<T> &v) {
  std::ofstream f(path, std::ios::binary);
  if (!v.empty()) f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
  return bool(f);
}

template <typename T>
T *to_device(const std::vector<T> &h) {
  const int64_t bytes = (int64_t)(h.size() ? h.size() : 1) * (int64_t)sizeof(T);
  T *d = static_cast<T *>(spx_malloc(bytes));
  if (d && !h.empty() && spx_memcpy_h2d(d, h.data(), (int64_t)(h.size() * sizeof(T))) != SPX_OK) {
    spx_free(d);
    return nullptr;
  }
  return d;
}

int host_checks() {
  // auto sigma on nnz/m (anonymouslib_cuda.h:297-313), m = 10
  const long long nnzs[] = {0, 30, 40, 41, 50, 320, 321, 330, 2560, 2569, 2570, 100000};
  for (long long z : nnzs) {
    anonymouslibHandle<int, unsigned int, double> A(10, 10);
    A.inputCSR((int)z, nullptr, nullptr, nullptr);
    A.setSigma(ANONYMOUSLIB_AUTO_TUNED_SIGMA);
    std::printf("{\"check\": \"auto_sigma\", \"m\": 10, \"nnz\": %lld, \"sigma\": %d}\n", z, A.sigma());
  }
  {
    anonymouslibHandle<int, unsigned int, double> A(0, 0);
    A.inputCSR(0, nullptr, nullptr, nullptr);
    A.setSigma(ANONYMOUSLIB_AUTO_TUNED_SIGMA);
    std::printf("{\"check\": \"auto_sigma\", \"m\": 0, \"nnz\": 0, \"sigma\": %d}\n", A.sigma());
  }
  {
    anonymouslibHandle<int, unsigned int, double> A(10, 10);
    A.inputCSR(30, nullptr, nullptr, nullptr);
    A.setSigma(12);
    const int s = A.sigma();
    double y = 0;
    const int spmv_csr = A.spmv(1.0, &y);  // still CSR: no device work
    A.setSigma(2);
    const int bad_lo = A.asCSR5();
    A.setSigma(33);
    const int bad_hi = A.asCSR5();
    std::printf("{\"check\": \"codes\", \"explicit_sigma\": %d, \"spmv_in_csr\": %d, \"sigma2\": %d, "
                "\"sigma33\": %d, \"format\": %d}\n",
                s, spmv_csr, bad_lo, bad_hi, A.format());
  }
  {
    anonymouslibHandle<int, unsigned int, float> F(10, 10);
    F.inputCSR(30, nullptr, nullptr, nullptr);
    float x = 0;
    std::printf("{\"check\": \"float\", \"setX\": %d, \"asCSR5\": %d}\n", F.setX(&x), F.asCSR5());
  }
  return 0;
}

int spmv_checks(const std::string &dir) {
  int m = 0, n = 0, nnz = 0;
  double a1 = 0, a2 = 0;
  int sig = 0;
  {
    FILE *f = std::fopen((dir + "/meta.txt").c_str(), "r");
    if (!f || std::fscanf(f, "%d %d %d %lf %lf %d", &m, &n, &nnz, &a1, &a2, &sig) != 6) {
      std::printf("{\"error\": \"meta.txt\"}\n");
      if (f) std::fclose(f);
      return 2;
    }
    std::fclose(f);
  }
  std::vector<int32_t> rp, col;
  std::vector<double> val, x1, x2;
  if (!read_bin(dir + "/rowptr.bin", rp, (size_t)m + 1) || !read_bin(dir + "/col.bin", col, (size_t)nnz) ||
      !read_bin(dir + "/val.bin", val, (size_t)nnz) || !read_bin(dir + "/x1.bin", x1, (size_t)n) ||
      !read_bin(dir + "/x2.bin", x2, (size_t)n)) {
    std::printf("{\"error\": \"inputs\"}\n");
    return 2;
  }
  int32_t *d_rp = to_device(rp), *d_col = to_device(col);
  double *d_val = to_device(val), *d_x1 = to_device(x1), *d_x2 = to_device(x2);
  double *d_y = static_cast<double *>(spx_malloc((int64_t)(m ? m : 1) * 8));
  if (!d_rp || !d_col || !d_val || !d_x1 || !d_x2 || !d_y) {
    std::printf("{\"error\": \"device memory\"}\n");
    return 2;
  }
  auto fetch_y = [&](const char *name) {
    std::vector<double> h((size_t)m);
    if (m) spx_memcpy_d2h(h.data(), d_y, (int64_t)m * 8);
    write_bin(dir + "/" + name, h);
  };
  // poison y: spmv overwrites it
  const std::vector<double> poison((size_t)m, 12345.0);
  if (m) spx_memcpy_h2d(d_y, poison.data(), (int64_t)m * 8);

  int rc[12];
  int k = 0;
  anonymouslibHandle<int, unsigned int, double> A(m, n);
  // the reference wrapper's order (csr5_kernel.cu:17-30)
  rc[k++] = A.inputCSR(nnz, d_rp, d_col, d_val);
  rc[k++] = A.setX(d_x1);
  A.setSigma(ANONYMOUSLIB_AUTO_TUNED_SIGMA);
  rc[k++] = A.warmup();
  rc[k++] = A.asCSR5();
  const int sigma_auto = A.sigma();
  rc[k++] = A.spmv(a1, d_y);
  spx_synchronize();
  fetch_y("y1.bin");
  // the caller's arrays are untouched by the conversion
  std::vector<int32_t> col_back((size_t)nnz);
  std::vector<double> val_back((size_t)nnz);
  if (nnz) {
    spx_memcpy_d2h(col_back.data(), d_col, (int64_t)nnz * 4);
    spx_memcpy_d2h(val_back.data(), d_val, (int64_t)nnz * 8);
  }
  const bool unchanged = col_back == col && std::memcmp(val_back.data(), val.data(), (size_t)nnz * 8) == 0;
  // a second x through the same conversion (the solver's loop)
  rc[k++] = A.setX(d_x2);
  rc[k++] = A.spmv(a2, d_y);
  spx_synchronize();
  fetch_y("y2.bin");
  // back to CSR, another sigma, convert again
  rc[k++] = A.asCSR();
  const int spmv_after_ascsr = A.spmv(a1, d_y);
  A.setSigma(sig);
  rc[k++] = A.setX(d_x1);
  rc[k++] = A.asCSR5();
  rc[k++] = A.spmv(a1, d_y);
  spx_synchronize();
  fetch_y("y3.bin");
  // the segmented engine through the same handle
  A.setKernel(SPX_KERNEL_SEG);
  A.asCSR();
  rc[k++] = A.asCSR5();
  A.setX(d_x2);
  const int rc_last = A.spmv(a2, d_y);
  spx_synchronize();
  fetch_y("y4.bin");
  const int rc_destroy = A.destroy();
  std::printf("{\"check\": \"spmv\", \"m\": %d, \"n\": %d, \"nnz\": %d, \"sigma_auto\": %d, \"sigma\": %d, "
              "\"unchanged\": %s, \"spmv_after_asCSR\": %d, \"destroy\": %d, \"format\": %d, \"rc\": [",
              m, n, nnz, sigma_auto, sig, unchanged ? "true" : "false", spmv_after_ascsr, rc_destroy, A.format());
  for (int i = 0; i < k; i++) std::printf("%d, ", rc[i]);
  std::printf("%d]}\n", rc_last);
  spx_free(d_rp); spx_free(d_col); spx_free(d_val); spx_free(d_x1); spx_free(d_x2); spx_free(d_y);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc >= 2 && std::strcmp(argv[1], "host") == 0) return host_checks();
  if (argc >= 3 && std::strcmp(argv[1], "spmv") == 0) return spmv_checks(argv[2]);
  std::fprintf(stderr, "usage: handle_check host | spmv <dir>\n");
  return 2;
}
