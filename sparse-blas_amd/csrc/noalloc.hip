// noalloc.hip -- y = alpha*A*x + beta*y on the caller's device CSR with no
// device allocation: the entry point for callers that cannot allocate
// (code under stream capture, tight serving loops, harnesses that forbid
// hipMalloc).  It computes what spx_dropin_csr5_kernel computes, on the
// same arguments, by the merge-path kernel (kernel 2) run directly on the
// caller's arrays.  Its only device memory beyond the inputs and outputs --
// the merge-path unit table and the units' carry slots -- lives in one
// file-scope __device__ array of kNoAllocScratch bytes per GPU.
//
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (other host threads, other streams) race and must be serialised by the
// caller.  Each call rewrites every entry it reads (all unit records and
// both carry slots of every unit), so nothing needs clearing between calls.
#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/spx_capi.h"
#include "spx_build.h"
#include "spx_device.h"

namespace spx {

constexpr int64_t kNoAllocScratch = 8 << 20;  // bytes per GPU
constexpr int kNoAllocItems = 8;              // merge items per thread
constexpr int kMaxDevices = 64;

// A 256-byte header (the input check's flag), then per unit: its record
// (32 bytes), 2 carry rows (4 bytes each) and 2 carry values (8 bytes each).
constexpr int64_t kNoAllocHeader = 256;
constexpr int64_t kNoAllocPerUnit = (int64_t)sizeof(MUnit) + 2 * sizeof(int32_t) + 2 * sizeof(double);
constexpr int64_t kNoAllocMaxUnits = (kNoAllocScratch - kNoAllocHeader) / kNoAllocPerUnit;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

namespace {

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

// The device address of this device's scratch array, taken once per device
// (with that device current).  Host code uses only this address: the
// array's name is a host-side stand-in.
int scratch_ptr(void **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_last_error("no-alloc SpMV: no current device");
    return SPX_EDEV;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_last_error(std::string("no-alloc SpMV: hipGetSymbolAddress: ") + hipGetErrorString(e));
      return SPX_EDEV;
    }
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return SPX_OK;
}

}  // namespace

// Merge-path coordinates of one diagonal of the whole matrix (one partition:
// rows [0, m), nonzeros [0, nnz)): the first row end not consumed and the
// first nonzero not consumed after d items.
__device__ inline void merge_coord(const int32_t *rowptr, int m, int nnz, int64_t d, int &row, int &nz) {
  int64_t lo = d - nnz > 0 ? d - nnz : 0, hi = d < m ? d : m;
  while (lo < hi) {
    const int64_t piv = (lo + hi) >> 1;
    if ((int64_t)rowptr[piv + 1] <= d - piv - 1)
      lo = piv + 1;
    else
      hi = piv;
  }
  row = (int)lo;
  nz = (int)(d - lo);
}

// *bad = 1 unless the row pointer runs monotonically from 0 to nnz and every
// column index lies in [0, n).
__global__ void k_noalloc_check(const int32_t *__restrict__ rowptr, int m, int nnz, const int32_t *__restrict__ col,
                                int n, int *__restrict__ bad) {
  int found = 0;
  const int64_t span = (int64_t)(m > nnz ? m : nnz);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < span; i += (int64_t)gridDim.x * blockDim.x) {
    if (i < m) found |= rowptr[i] > rowptr[i + 1];
    if (i < nnz) found |= (col[i] < 0) | (col[i] >= n);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) found |= (rowptr[0] != 0) | (rowptr[m] != nnz);
  if (found) atomicOr(bad, 1);
}

// Unit u covers diagonals [u*tw, (u+1)*tw); its window fields receive its
// nonzero range for k_unit_window.
__global__ void k_noalloc_units(const int32_t *__restrict__ rowptr, int m, int nnz, int tw, int64_t units,
                                MUnit *__restrict__ out) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= units) return;
  const int64_t L = (int64_t)m + nnz;
  const int64_t d0 = u * tw < L ? u * tw : L, d1 = (u + 1) * tw < L ? (u + 1) * tw : L;
  int r0, j0, r1, j1;
  merge_coord(rowptr, m, nnz, d0, r0, j0);
  merge_coord(rowptr, m, nnz, d1, r1, j1);
  out[u] = MUnit{r0, j0, r1, j1, j0, j1, 0, 0};
}

}  // namespace spx

using namespace spx;

extern "C" int spx_dropin_csr5_kernel_noalloc(int m, int n, int nnz, double *alpha, double *csrVal, int *csrRowPtr,
                                              int *csrColIndex, double *x, double *beta, double *y) {
  // the same arguments as spx_dropin_csr5_kernel (device pointers, int32 CSR)
  if (m < 0 || n < 0 || nnz < 0 || !alpha || !beta) {
    set_last_error("no-alloc SpMV: bad sizes or null alpha/beta");
    return -1;
  }
  if (m == 0) return 0;
  const int tw = kBlock * kNoAllocItems;
  const int64_t units = ((int64_t)m + nnz + tw - 1) / tw;
  if (units > kNoAllocMaxUnits) {
    set_last_error("no-alloc SpMV: m + nnz = " + std::to_string((int64_t)m + nnz) + " exceeds the " +
                   std::to_string(kNoAllocMaxUnits * tw) + " that its " + std::to_string(kNoAllocScratch >> 20) +
                   " MiB scratch array holds");
    return -1;
  }
  if (!csrRowPtr || !y || (nnz > 0 && (!csrVal || !csrColIndex || !x))) {
    set_last_error("no-alloc SpMV: null matrix or vector");
    return -1;
  }
  void *base = nullptr;
  if (scratch_ptr(&base) != SPX_OK) return -1;
  const double a = *alpha, b = *beta;
  hipStream_t st = nullptr;
  // the input check: its flag is cleared by a kernel (no memset node, so a
  // captured call stays replayable) and read back before any SpMV kernel
  int *bad = static_cast<int *>(base);
  hipError_t e = launch_fill_i32(bad, 0, 1, st);
  if (e == hipSuccess) {
    const int64_t span = std::max<int64_t>(m, nnz);
    const unsigned g = (unsigned)std::min<int64_t>((span + 255) / 256, 4096);
    hipLaunchKernelGGL(k_noalloc_check, dim3(g), dim3(256), 0, st, csrRowPtr, m, nnz, csrColIndex, n, bad);
    e = hipGetLastError();
  }
  int hbad = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    set_last_error(std::string("no-alloc SpMV: ") + hipGetErrorString(e));
    return -1;
  }
  if (hbad) {
    set_last_error("no-alloc SpMV: row pointer not monotone from 0 to nnz, or a column index outside [0, n)");
    return -1;
  }
  if (nnz == 0) {
    e = launch_scale_y(b, y, m, st);
  } else {
    unsigned char *u0 = static_cast<unsigned char *>(base) + kNoAllocHeader;
    MUnit *mu = reinterpret_cast<MUnit *>(u0);
    int32_t *crow = reinterpret_cast<int32_t *>(u0 + units * sizeof(MUnit));
    double *cval = reinterpret_cast<double *>(u0 + units * sizeof(MUnit) + 2 * units * sizeof(int32_t));
    hipLaunchKernelGGL(k_noalloc_units, dim3((unsigned)((units + 255) / 256)), dim3(256), 0, st, csrRowPtr, m, nnz,
                       tw, units, mu);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_unit_window(csrColIndex, reinterpret_cast<int32_t *>(mu), 8, 4, units, st);
    if (e == hipSuccess) {
      KArgs<double> k = {};
      k.unit_begin = 0;
      k.unit_count = units;
      k.rowptr = csrRowPtr;
      k.col = csrColIndex;
      k.val = csrVal;
      k.x = x;
      k.alpha = a;
      k.beta = b;
      k.y = y;
      k.crow = crow;
      k.cval = cval;
      k.munits = mu;
      e = launch_merge_f64(k, kNoAllocItems, direct_grid(units), st);
    }
    if (e == hipSuccess)
      e = launch_finish(nullptr, 0, nullptr, nullptr, nullptr, false, crow, cval, 2 * units, a, b, y, st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    set_last_error(std::string("no-alloc SpMV: ") + hipGetErrorString(e));
    return -1;
  }
  return 0;
}
