// dropin.cpp -- the reference's entry points (SparseBLAS/include/spmv_kernel.h)
// on one MI355X.  Each call is stateless like the reference's: it builds a
// device plan from the caller's host arrays, copies x (and y when beta != 0)
// in, executes, copies y out and frees everything.  The copies go through a
// pinned double buffer (stage.cpp), and the matrix's copy overlaps the
// host's pass over the row pointer.  Return codes follow the
// reference: 0 ok, -1 bad argument / memory guard / compute failure, 1 device
// allocation or copy failure.
#include <hip/hip_runtime.h>
#include <sys/time.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "../../include/spmv_kernel.h"
#include "../../include/spx_capi.h"
#include "spx_internal.h"

namespace {

std::atomic<long long> g_free_override{-1};
std::atomic<int> g_v1_streams{1};
thread_local double g_times[7] = {0, 0, 0, 0, 0, 0, 0};

double ms_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

// Footprint of one partition as the reference's guard counts it
// (SB/src/dspmv_mgpu_v1.cu:107-112): values, columns, int32 row pointer, x, y.
double footprint_gb(long long nnz, long long rows, int n) {
  long long b = nnz * (long long)sizeof(double) + nnz * (long long)sizeof(int) +
                (rows + 1) * (long long)sizeof(int) + (long long)n * sizeof(double) +
                rows * (long long)sizeof(double);
  return (double)b / 1e9;
}

int code_of(int rc) {
  if (rc == SPX_OK) return 0;
  if (rc == SPX_EMEM) return 1;
  return -1;
}

// Host arrays in, host y out; one plan.
int run_host(const spx_plan_desc &d, double alpha, const double *x, double beta, double *y) {
  const auto t_all = std::chrono::steady_clock::now();
  for (double &t : g_times) t = 0.0;
  spx_plan_t plan = nullptr;
  auto t = std::chrono::steady_clock::now();
  int rc = spx_plan_create(&plan, &d);
  g_times[0] = ms_since(t);
  if (rc != SPX_OK) {
    std::fprintf(stderr, "spx: plan creation failed: %s\n", spx_last_error());
    return code_of(rc);
  }
  int64_t info[24] = {0};
  if (spx_plan_info(plan, info, 24) >= 23) g_times[1] = (double)info[22] / 1000.0;
  double *dx = nullptr, *dy = nullptr;
  int ret = 0;
  if (d.m > 0) {
    t = std::chrono::steady_clock::now();
    dx = (double *)spx_malloc((int64_t)std::max(d.n, 1) * 8);
    dy = (double *)spx_malloc((int64_t)d.m * 8);
    if (!dx || !dy) {
      ret = 1;
    } else if ((d.n > 0 && spx::stage_h2d(dx, x, (size_t)d.n * 8) != hipSuccess) ||
               (beta != 0.0 && spx::stage_h2d(dy, y, (size_t)d.m * 8) != hipSuccess)) {
      ret = 1;
    }
    g_times[2] = ms_since(t);
    t = std::chrono::steady_clock::now();
    if (ret == 0 &&
        (spx_plan_execute(plan, alpha, dx, beta, dy, nullptr) != SPX_OK || spx_synchronize() != SPX_OK)) {
      std::fprintf(stderr, "spx: execute failed: %s\n", spx_last_error());
      ret = -1;
    }
    g_times[3] = ms_since(t);
    t = std::chrono::steady_clock::now();
    if (ret == 0 && spx::stage_d2h(y, dy, (size_t)d.m * 8) != hipSuccess) ret = 1;
    g_times[4] = ms_since(t);
  }
  t = std::chrono::steady_clock::now();
  if (dx) spx_free(dx);
  if (dy) spx_free(dy);
  spx_plan_destroy(plan);
  g_times[5] = ms_since(t);
  g_times[6] = ms_since(t_all);
  return ret;
}

// spMV_mgpu_v2 for a matrix whose footprint exceeds the memory guard: its
// tasks (SB/src/dspmv_mgpu_v2.cu:218-267: T = ceil(nnz / nb) nearly equal
// nonzero ranges, start = floor(t nnz / T)) run one after another through
// plans that each hold one task's nonzeros, so no more than nb of them are
// on the device at once (v2:108-115 allocates nb-sized buffers per stream
// and streams the tasks through them).  x is uploaded once; each task's row
// sums (alpha applied) come back and are added on the host in task order,
// and beta * y is added last (v2:381-437 gathers the tasks' results on the
// host likewise).
int run_streamed(int m, int n, long long nnz, double alpha, const double *val, const long long *rowptr,
                 const int *col, const double *x, double beta, double *y, int kernel, long long nb) {
  const auto t_all = std::chrono::steady_clock::now();
  for (double &t : g_times) t = 0.0;
  // the tasks' row ranges come from the row pointer: it must run
  // monotonically from 0 to nnz
  if (m > 0 && (!rowptr || rowptr[0] != 0 || rowptr[m] != nnz)) return -1;
  for (int r = 0; r < m; r++)
    if (rowptr[r + 1] < rowptr[r]) return -1;
  if (nnz > 0 && (!val || !col)) return -1;
  const long long T = (nnz + nb - 1) / nb;
  std::vector<double> acc((size_t)m, 0.0);
  double *dx = (double *)spx_malloc((int64_t)std::max(n, 1) * 8);
  double *dy = nullptr;
  int ret = 0;
  if (!dx || (n > 0 && spx::stage_h2d(dx, x, (size_t)n * 8) != hipSuccess)) ret = 1;
  std::vector<long long> lrp;
  std::vector<double> part;
  int64_t dy_rows = 0;
  for (long long t = 0; t < T && ret == 0; t++) {
    const long long s0 = (long long)std::floor((double)(t * nnz / T));
    const long long e1 = (long long)std::floor((double)((t + 1) * nnz / T));  // one past the task's last nonzero
    if (e1 <= s0) continue;
    // rows holding the task's first and last nonzeros: upper_bound - 1
    const int r0 = (int)(std::upper_bound(rowptr, rowptr + m + 1, s0) - rowptr) - 1;
    const int r1 = (int)(std::upper_bound(rowptr, rowptr + m + 1, e1 - 1) - rowptr) - 1;
    const int mt = r1 - r0 + 1;
    lrp.assign((size_t)mt + 1, 0);
    for (int i = 0; i <= mt; i++)
      lrp[(size_t)i] = std::min<long long>(std::max<long long>(rowptr[r0 + i] - s0, 0), e1 - s0);
    if (mt > dy_rows) {
      if (dy) spx_free(dy);
      dy_rows = mt;
      dy = (double *)spx_malloc(dy_rows * 8);
      if (!dy) {
        ret = 1;
        break;
      }
    }
    spx_plan_desc d;
    spx_plan_desc_init(&d);
    d.m = mt; d.n = n; d.nnz = e1 - s0;
    d.rowptr64 = reinterpret_cast<const int64_t *>(lrp.data());
    d.col = col + s0; d.val = val + s0;
    d.kernel = kernel;
    d.mode = SPX_MODE_NNZ;
    d.parts = 1;
    spx_plan_t plan = nullptr;
    int rc = spx_plan_create(&plan, &d);
    if (rc == SPX_OK) {
      rc = spx_plan_execute(plan, alpha, dx, 0.0, dy, nullptr);
      if (rc == SPX_OK) rc = spx_synchronize();
      spx_plan_destroy(plan);
    }
    if (rc != SPX_OK) {
      std::fprintf(stderr, "spx: streamed task %lld failed: %s\n", t, spx_last_error());
      ret = code_of(rc);
      break;
    }
    part.resize((size_t)mt);
    if (spx::stage_d2h(part.data(), dy, (size_t)mt * 8) != hipSuccess) {
      ret = 1;
      break;
    }
    for (int i = 0; i < mt; i++) acc[(size_t)(r0 + i)] += part[(size_t)i];
  }
  if (dx) spx_free(dx);
  if (dy) spx_free(dy);
  if (ret == 0)
    for (int r = 0; r < m; r++) y[r] = beta == 0.0 ? acc[(size_t)r] : acc[(size_t)r] + beta * y[r];
  g_times[6] = ms_since(t_all);
  return ret;
}

bool common_args_ok(int m, int n, long long nnz, double *alpha, double *beta, double *y, int ngpu) {
  if (ngpu != 1) {
    std::fprintf(stderr, "spx: ngpu must be 1 (this library drives one MI355X)\n");
    return false;
  }
  if (m < 0 || n < 0 || nnz < 0 || !alpha || !beta || (m > 0 && !y)) return false;
  return true;
}

}  // namespace

extern "C" {

void spx_debug_set_free_mem(long long bytes) { g_free_override = bytes; }

int spx_dropin_last_times(double *ms, int cap) {
  if (!ms || cap <= 0) return 0;
  const int c = std::min(cap, 7);
  for (int i = 0; i < c; i++) ms[i] = g_times[i];
  return c;
}

int spx_set_v1_streams(int parts) {
  int old = g_v1_streams.load();
  if (parts >= 1) g_v1_streams = parts;
  return old;
}

int spx_get_row_from_index(int n, long long *a, long long idx) {
  // SB/src/spmv_helper.cu:16-39, behaviour kept exactly (including the
  // early return on the first probe equal to idx)
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
  if (idx == a[l]) return l;
  if (idx == a[r]) return r;
  return l;
}

double spx_get_time(void) {
  // SB/src/spmv_helper.cu:41-48
  struct timeval tp;
  gettimeofday(&tp, nullptr);
  return (double)tp.tv_sec + (double)tp.tv_usec / 1e6;
}

double spx_get_gpu_availble_mem(int ngpu) {
  // SB/src/spmv_helper.cu:51-76: minimum free memory (GB) over the GPUs
  long long ov = g_free_override.load();
  if (ov >= 0) return (double)ov / 1e9;
  double mn = std::numeric_limits<double>::max();
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (int d = 0; d < ngpu; d++) {
    size_t fr = 0, tot = 0;
    if (hipSetDevice(d) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) {
      mn = 0;
      break;
    }
    mn = std::min(mn, (double)fr / 1e9);
  }
  (void)hipSetDevice(cur);
  return mn;
}

int spx_dropin_baseline(int m, int n, long long nnz, double *alpha, double *csrVal, long long *csrRowPtr,
                        int *csrColIndex, double *x, double *beta, double *y, int ngpu) {
  if (!common_args_ok(m, n, nnz, alpha, beta, y, ngpu)) return -1;
  // SB/src/dspmv_mgpu_baseline.cu:70-79
  if (footprint_gb(nnz, m, n) > 0.8 * spx_get_gpu_availble_mem(ngpu)) return -1;
  spx_plan_desc d;
  spx_plan_desc_init(&d);
  d.m = m; d.n = n; d.nnz = nnz;
  d.rowptr64 = reinterpret_cast<const int64_t *>(csrRowPtr);
  d.col = csrColIndex; d.val = csrVal;
  d.kernel = SPX_KERNEL_VECTOR;
  d.mode = SPX_MODE_ROWS;
  d.parts = 1;
  return run_host(d, *alpha, x, *beta, y);
}

int spx_dropin_v1(int m, int n, long long nnz, double *alpha, double *csrVal, long long *csrRowPtr,
                  int *csrColIndex, double *x, double *beta, double *y, int ngpu, int kernel) {
  if (!common_args_ok(m, n, nnz, alpha, beta, y, ngpu)) return -1;
  if (kernel < 1 || kernel > 3) return -1;
  // SB/src/dspmv_mgpu_v1.cu:106-117 (one partition per GPU at ngpu = 1)
  if (footprint_gb(nnz, m, n) > 0.8 * spx_get_gpu_availble_mem(ngpu)) return -1;
  spx_plan_desc d;
  spx_plan_desc_init(&d);
  d.m = m; d.n = n; d.nnz = nnz;
  d.rowptr64 = reinterpret_cast<const int64_t *>(csrRowPtr);
  d.col = csrColIndex; d.val = csrVal;
  d.kernel = kernel;
  d.mode = SPX_MODE_NNZ;
  d.parts = g_v1_streams.load();
  return run_host(d, *alpha, x, *beta, y);
}

int spx_dropin_v2(int m, int n, long long nnz, double *alpha, double *csrVal, long long *csrRowPtr,
                  int *csrColIndex, double *x, double *beta, double *y, int ngpu, int kernel, long long nb,
                  int copy_of_workspace) {
  if (!common_args_ok(m, n, nnz, alpha, beta, y, ngpu)) return -1;
  if (kernel < 1 || kernel > 3) return -1;
  const int q = copy_of_workspace;
  // SB/src/dspmv_mgpu_v2.cu:43-46
  if (q <= 0 || nb <= 0) return -1;
  long long cap = (long long)(0.8 * spx_get_gpu_availble_mem(ngpu) * 1e9 /
                              (double)(sizeof(double) + sizeof(int) + sizeof(int))) / q;
  nb = std::min(nb, cap);
  if (nb <= 0) return -1;
  // a matrix beyond the guard streams through nb-sized tasks (v2:108-115)
  if (footprint_gb(nnz, m, n) > 0.8 * spx_get_gpu_availble_mem(ngpu))
    return run_streamed(m, n, nnz, *alpha, csrVal, csrRowPtr, csrColIndex, x, *beta, y, kernel, nb);
  spx_plan_desc d;
  spx_plan_desc_init(&d);
  d.m = m; d.n = n; d.nnz = nnz;
  d.rowptr64 = reinterpret_cast<const int64_t *>(csrRowPtr);
  d.col = csrColIndex; d.val = csrVal;
  d.kernel = kernel;
  d.mode = SPX_MODE_QUEUE;
  d.nb = nb;
  d.queues = q;
  return run_host(d, *alpha, x, *beta, y);
}

int spx_dropin_csr5_kernel(int m, int n, int nnz, double *alpha, double *csrVal, int *csrRowPtr, int *csrColIndex,
                           double *x, double *beta, double *y) {
  // SB/src/csr5_kernel.cu:12-32 on device pointers; alpha and beta honoured
  if (m < 0 || n < 0 || nnz < 0 || !alpha || !beta) return -1;
  spx_plan_desc d;
  spx_plan_desc_init(&d);
  d.m = m; d.n = n; d.nnz = nnz;
  d.rowptr32 = csrRowPtr;
  d.col = csrColIndex; d.val = csrVal;
  d.kernel = SPX_KERNEL_TILE;
  d.mode = SPX_MODE_NNZ;
  d.parts = 1;
  d.flags = SPX_FLAG_DEVICE_PTRS;
  spx_plan_t plan = nullptr;
  int rc = spx_plan_create(&plan, &d);
  if (rc != SPX_OK) return code_of(rc);
  rc = spx_plan_execute(plan, *alpha, x, *beta, y, nullptr);
  if (rc == SPX_OK) rc = spx_synchronize();
  spx_plan_destroy(plan);
  return rc == SPX_OK ? 0 : -1;
}

}  // extern "C"

// ---- C++ linkage, as declared by the reference header ---------------------------

int csr5_kernel(int m, int n, int nnz, double *alpha, double *csrVal, int *csrRowPtr, int *csrColIndex, double *x,
                double *beta, double *y) {
  return spx_dropin_csr5_kernel(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y);
}

int spMV_mgpu_baseline(int m, int n, long long nnz, double *alpha, double *csrVal, long long *csrRowPtr,
                       int *csrColIndex, double *x, double *beta, double *y, int ngpu) {
  return spx_dropin_baseline(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, ngpu);
}

int spMV_mgpu_v1(int m, int n, long long nnz, double *alpha, double *csrVal, long long *csrRowPtr, int *csrColIndex,
                 double *x, double *beta, double *y, int ngpu, int kernel) {
  return spx_dropin_v1(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, ngpu, kernel);
}

int spMV_mgpu_v2(int m, int n, long long nnz, double *alpha, double *csrVal, long long *csrRowPtr, int *csrColIndex,
                 double *x, double *beta, double *y, int ngpu, int kernel, long long nb, int copy_of_workspace) {
  return spx_dropin_v2(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, ngpu, kernel, nb,
                       copy_of_workspace);
}

int get_row_from_index(int n, long long *a, long long idx) { return spx_get_row_from_index(n, a, idx); }

double get_time() { return spx_get_time(); }

double get_gpu_availble_mem(int ngpu) { return spx_get_gpu_availble_mem(ngpu); }
