/*
 * spmv_kernel.h -- drop-in C++ API of the MI355X (gfx950) FP64 CSR SpMV
 * library libspx.so.  The declarations match the reference header
 * SparseBLAS/include/spmv_kernel.h:4-34 name for name and argument for
 * argument, so a caller of the reference recompiles against this header and
 * links -lspx instead of the CUDA objects.
 *
 * All calls compute y = alpha*A*x + beta*y for an m-by-n CSR matrix A
 * (0-based, general).  Pointers are HOST memory except in csr5_kernel.
 * Return codes follow the reference: 0 success; -1 a bad argument (ngpu != 1,
 * kernel not in {1,2,3}, nb <= 0, copy_of_workspace <= 0), a footprint above
 * 0.8 of free device memory (baseline and v1; v2 then streams its tasks), or
 * a failed computation; 1 a device allocation or copy failure.  Each entry point is a thin wrapper over the extern "C"
 * ABI in spx_capi.h.
 *
 * Differences from the reference (all deliberate):
 *  - ngpu must be 1 (one MI355X); the nnz-balanced split of v1/v2 runs over
 *    HIP streams and workgroups of that GPU instead of over GPUs.
 *  - kernel 3 honours alpha and beta (the reference's CSR5 path drops both,
 *    SparseBLAS/include/detail/cuda/csr5_spmv_cuda.h:22 and
 *    SparseBLAS/src/csr5_kernel.cu:30).
 *  - spMV_mgpu_v2 returns 0 on success (the reference falls off the end,
 *    SparseBLAS/src/dspmv_mgpu_v2.cu:207).
 *  - results are bitwise reproducible: split rows are merged in a fixed
 *    order on the device, never by atomics or completion order.
 */
#ifndef SPMV_KERNEL
#define SPMV_KERNEL

/* SparseBLAS/include/spmv_kernel.h:4-7: kernel version 3 on one partition
 * held in DEVICE memory (int32 row pointer). */
int csr5_kernel(int m, int n, int nnz, double *alpha,
                double *csrVal, int *csrRowPtr, int *csrColIndex,
                double *x, double *beta,
                double *y);

/* SparseBLAS/include/spmv_kernel.h:9-13: row-block split, kernel version 1. */
int spMV_mgpu_baseline(int m, int n, long long nnz, double *alpha,
                       double *csrVal, long long *csrRowPtr, int *csrColIndex,
                       double *x, double *beta,
                       double *y,
                       int ngpu);

/* SparseBLAS/include/spmv_kernel.h:14-19: nnz-balanced split, kernel 1/2/3. */
int spMV_mgpu_v1(int m, int n, long long nnz, double *alpha,
                 double *csrVal, long long *csrRowPtr, int *csrColIndex,
                 double *x, double *beta,
                 double *y,
                 int ngpu,
                 int kernel);

/* SparseBLAS/include/spmv_kernel.h:21-28: task pool of nb-nonzero tasks over
 * copy_of_workspace streams; here a device work queue drained by persistent
 * workgroups.  nb is capped at 0.8 of free memory / 16 bytes / copies
 * (SparseBLAS/src/dspmv_mgpu_v2.cu:43); a matrix whose footprint exceeds
 * 0.8 of free memory runs its tasks one after another, at most nb nonzeros
 * on the device at a time (v2:108-115). */
int spMV_mgpu_v2(int m, int n, long long nnz, double *alpha,
                 double *csrVal, long long *csrRowPtr, int *csrColIndex,
                 double *x, double *beta,
                 double *y,
                 int ngpu,
                 int kernel,
                 long long nb,
                 int copy_of_workspace);

/* SparseBLAS/include/spmv_kernel.h:30: binary search of a row pointer
 * a[0..n] for nonzero index idx, with the reference's behaviour kept
 * (on runs of equal entries it returns the first matching probe). */
int get_row_from_index(int n, long long *a, long long idx);

/* SparseBLAS/include/spmv_kernel.h:32: wall clock in seconds. */
double get_time();

/* SparseBLAS/include/spmv_kernel.h:34: free device memory in GB (minimum
 * over the first ngpu devices). */
double get_gpu_availble_mem(int ngpu);

#endif /* SPMV_KERNEL */
