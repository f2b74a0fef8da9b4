/*
 * anonymouslib_hip.h -- persistent SpMV handle for libspx.so with the
 * interface of the reference's CSR5 handle class
 * (SparseBLAS/include/anonymouslib_cuda.h:11-51, status and format codes of
 * SparseBLAS/include/detail/common.h:13-22): convert a device CSR matrix
 * once, then run y = alpha*A*x as often as an iterative solver needs.
 *
 *   anonymouslibHandle<int, unsigned int, double> A(m, n);
 *   A.inputCSR(nnz, d_rowptr, d_col, d_val);   // DEVICE pointers, int32
 *   A.setX(d_x);
 *   A.setSigma(ANONYMOUSLIB_AUTO_TUNED_SIGMA);
 *   A.warmup();
 *   A.asCSR5();                                 // the one-time conversion
 *   A.spmv(alpha, d_y);                         // many times
 *   A.destroy();
 *
 * Header-only C++ over the extern "C" plan layer of spx_capi.h: asCSR5
 * builds a plan (kernel 3, the CSR5-style tiles, with the handle's sigma;
 * setKernel picks another engine), spmv enqueues one execute of it on the
 * null stream, asCSR and destroy release it.
 *
 * Differences from the reference (all deliberate):
 *  - asCSR5 leaves the caller's arrays untouched: the plan keeps its own
 *    device copy in the tiled layout, so asCSR has nothing to transpose back
 *    (the reference transposes col/val in place, anonymouslib_cuda.h:78-103,
 *    225-231) and the caller may reuse or free its arrays after asCSR5.
 *  - spmv computes y = alpha*A*x and overwrites y (the reference drops alpha,
 *    SparseBLAS/include/detail/cuda/csr5_spmv_cuda.h:22, and adds the split
 *    rows' partial sums into y, so it needs y zeroed beforehand).
 *  - sigma is limited to 4..32 (the tile kernel's range); setSigma with
 *    another value makes asCSR5 return ANONYMOUSLIB_UNSUPPORTED_CSR5_OMEGA,
 *    the reference's code for tile parameters it cannot encode (:123-124).
 *  - ANONYMOUSLIB_CSR5_OMEGA is 64: a tile is one wave64 (the reference's is
 *    a 32-lane warp, common_cuda.h:11).
 *  - the index type must be 32-bit and the value type double (float values
 *    give ANONYMOUSLIB_UNSUPPORTED_VALUE_TYPE from setX and asCSR5).
 *  - the destructor releases the plan (the reference leaks unless destroy is
 *    called); handles are move-only.
 *  - sigma defaults to auto (the reference leaves it uninitialised until
 *    setSigma).
 */
#ifndef ANONYMOUSLIB_HIP_H
#define ANONYMOUSLIB_HIP_H

#include <stdint.h>

#include "spx_capi.h"

#define ANONYMOUSLIB_SUCCESS 0
#define ANONYMOUSLIB_UNKOWN_FORMAT -1
#define ANONYMOUSLIB_UNSUPPORTED_CSR5_OMEGA -2
#define ANONYMOUSLIB_CSR_TO_CSR5_FAILED -3
#define ANONYMOUSLIB_UNSUPPORTED_CSR_SPMV -4
#define ANONYMOUSLIB_UNSUPPORTED_VALUE_TYPE -5

#define ANONYMOUSLIB_FORMAT_CSR 0
#define ANONYMOUSLIB_FORMAT_CSR5 1

#define ANONYMOUSLIB_CSR5_OMEGA 64
#define ANONYMOUSLIB_AUTO_TUNED_SIGMA -1

template <class ANONYMOUSLIB_IT, class ANONYMOUSLIB_UIT, class ANONYMOUSLIB_VT>
class anonymouslibHandle {
 public:
  anonymouslibHandle(ANONYMOUSLIB_IT m, ANONYMOUSLIB_IT n) : _m(m), _n(n) {}
  ~anonymouslibHandle() { release(); }
  anonymouslibHandle(const anonymouslibHandle &) = delete;
  anonymouslibHandle &operator=(const anonymouslibHandle &) = delete;
  anonymouslibHandle(anonymouslibHandle &&o) noexcept { *this = static_cast<anonymouslibHandle &&>(o); }
  anonymouslibHandle &operator=(anonymouslibHandle &&o) noexcept {
    if (this != &o) {
      release();
      _format = o._format; _m = o._m; _n = o._n; _nnz = o._nnz;
      _csr_row_pointer = o._csr_row_pointer; _csr_column_index = o._csr_column_index; _csr_value = o._csr_value;
      _csr5_sigma = o._csr5_sigma; _sigma_request = o._sigma_request; _kernel = o._kernel;
      _plan = o._plan; _x = o._x;
      o._plan = nullptr;
      o._format = ANONYMOUSLIB_FORMAT_CSR;
    }
    return *this;
  }

  /* anonymouslib_cuda.h:54-58: bring the device up before timed work. */
  int warmup() { return spx_synchronize() == SPX_OK ? ANONYMOUSLIB_SUCCESS : ANONYMOUSLIB_UNKOWN_FORMAT; }

  /* anonymouslib_cuda.h:60-75: remember the device CSR arrays (m+1 row
   * pointers, nnz columns and values); the handle is in CSR format after. */
  int inputCSR(ANONYMOUSLIB_IT nnz, ANONYMOUSLIB_IT *csr_row_pointer, ANONYMOUSLIB_IT *csr_column_index,
               ANONYMOUSLIB_VT *csr_value) {
    release();
    _format = ANONYMOUSLIB_FORMAT_CSR;
    _nnz = nnz;
    _csr_row_pointer = csr_row_pointer;
    _csr_column_index = csr_column_index;
    _csr_value = csr_value;
    return ANONYMOUSLIB_SUCCESS;
  }

  /* anonymouslib_cuda.h:77-103: back to CSR; releases the tiled copy. */
  int asCSR() {
    release();
    _format = ANONYMOUSLIB_FORMAT_CSR;
    return ANONYMOUSLIB_SUCCESS;
  }

  /* anonymouslib_cuda.h:105-236: the one-time conversion to the tiled
   * format (a plan of the chosen engine, kernel 3 by default). */
  int asCSR5() {
    if (_format == ANONYMOUSLIB_FORMAT_CSR5) return ANONYMOUSLIB_SUCCESS;
    if (sizeof(ANONYMOUSLIB_VT) != sizeof(double) || sizeof(ANONYMOUSLIB_IT) != sizeof(int32_t))
      return ANONYMOUSLIB_UNSUPPORTED_VALUE_TYPE;
    _csr5_sigma = computeSigma();
    if (_kernel == SPX_KERNEL_TILE && (_csr5_sigma < 4 || _csr5_sigma > 32))
      return ANONYMOUSLIB_UNSUPPORTED_CSR5_OMEGA;
    spx_plan_desc d;
    spx_plan_desc_init(&d);
    d.m = (int)_m;
    d.n = (int)_n;
    d.nnz = (int64_t)_nnz;
    d.rowptr32 = reinterpret_cast<const int32_t *>(_csr_row_pointer);
    d.col = reinterpret_cast<const int32_t *>(_csr_column_index);
    d.val = reinterpret_cast<const double *>(_csr_value);
    d.kernel = _kernel;
    d.mode = SPX_MODE_NNZ;
    d.parts = 1;
    d.sigma = _kernel == SPX_KERNEL_TILE ? _csr5_sigma : 0;
    d.flags = SPX_FLAG_DEVICE_PTRS;
    spx_plan_t p = nullptr;
    if (spx_plan_create(&p, &d) != SPX_OK) return ANONYMOUSLIB_CSR_TO_CSR5_FAILED;
    _plan = p;
    _format = ANONYMOUSLIB_FORMAT_CSR5;
    return ANONYMOUSLIB_SUCCESS;
  }

  /* anonymouslib_cuda.h:238-272: the device x (n entries) later spmv calls
   * read. */
  int setX(ANONYMOUSLIB_VT *x) {
    _x = x;
    return sizeof(ANONYMOUSLIB_VT) == sizeof(double) ? ANONYMOUSLIB_SUCCESS : ANONYMOUSLIB_UNSUPPORTED_VALUE_TYPE;
  }

  /* anonymouslib_cuda.h:274-293: y = alpha*A*x (device y, m entries),
   * enqueued on the null stream; CSR format cannot run it. */
  int spmv(const ANONYMOUSLIB_VT alpha, ANONYMOUSLIB_VT *y) {
    if (_format != ANONYMOUSLIB_FORMAT_CSR5 || !_plan) return ANONYMOUSLIB_UNSUPPORTED_CSR_SPMV;
    return spx_plan_execute(_plan, (double)alpha, reinterpret_cast<const double *>(_x), 0.0,
                            reinterpret_cast<double *>(y), nullptr) == SPX_OK
               ? ANONYMOUSLIB_SUCCESS
               : ANONYMOUSLIB_UNKOWN_FORMAT;
  }

  /* anonymouslib_cuda.h:295-300 */
  int destroy() { return asCSR(); }

  /* anonymouslib_cuda.h:302-327: sigma, or the auto rule on nnz/m
   * (<= 4 -> 4; 5..32 -> itself; 33..256 -> 32; above -> 6). */
  void setSigma(int sigma) {
    _sigma_request = sigma;
    _csr5_sigma = computeSigma();
  }

  /* Additions.  The engine asCSR5 builds (SPX_KERNEL_*; kernel 4, the
   * segmented engine, is the fastest on skewed matrices); takes effect at
   * the next asCSR5. */
  void setKernel(int kernel) { _kernel = kernel; }
  int sigma() const { return _csr5_sigma; }
  int format() const { return _format; }
  spx_plan_t plan() const { return _plan; }

 private:
  int computeSigma() const {
    if (_sigma_request != ANONYMOUSLIB_AUTO_TUNED_SIGMA) return _sigma_request;
    const long long per_row = _m > 0 ? (long long)_nnz / (long long)_m : 0;
    if (per_row <= 4) return 4;
    if (per_row <= 32) return (int)per_row;
    if (per_row <= 256) return 32;
    return 6;
  }
  void release() {
    if (_plan) spx_plan_destroy(_plan);
    _plan = nullptr;
  }

  int _format = ANONYMOUSLIB_FORMAT_CSR;
  ANONYMOUSLIB_IT _m = 0;
  ANONYMOUSLIB_IT _n = 0;
  ANONYMOUSLIB_IT _nnz = 0;
  ANONYMOUSLIB_IT *_csr_row_pointer = nullptr;
  ANONYMOUSLIB_IT *_csr_column_index = nullptr;
  ANONYMOUSLIB_VT *_csr_value = nullptr;
  int _csr5_sigma = 4;
  int _sigma_request = ANONYMOUSLIB_AUTO_TUNED_SIGMA;
  int _kernel = SPX_KERNEL_TILE;
  spx_plan_t _plan = nullptr;
  ANONYMOUSLIB_VT *_x = nullptr;
};

#endif /* ANONYMOUSLIB_HIP_H */
