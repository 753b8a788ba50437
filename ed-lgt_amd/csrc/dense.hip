// Every eigenpair of a stored sector matrix above the Lanczos full-basis
// limit: the CSR entries scattered into a dense column-major N x N matrix on
// the device, then rocSOLVER's divide-and-conquer Hermitian (zheevd) or
// symmetric (dsyevd) eigensolver in place.  The reference's n_eigs="full"
// takes the same dense route on the host (qmb_hamiltonian.py:293-301: H as a
// dense array, scipy.linalg.eigh).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "common.h"
#include "kernels.h"

namespace edlgt {
namespace {

constexpr int kTB = 256;

// One thread per row: its entries land in column-major A (A[c * n + r]); the
// canonical CSR holds each (r, c) once, so the stores never collide.
template <class T>
__global__ void __launch_bounds__(kTB) csr_to_dense_colmajor(const int64_t* __restrict__ row_ptr,
                                                             const int32_t* __restrict__ col,
                                                             const T* __restrict__ val, int64_t n,
                                                             T* __restrict__ A) {
  for (int64_t r = (int64_t)blockIdx.x * kTB + threadIdx.x; r < n; r += (int64_t)gridDim.x * kTB) {
    for (int64_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) A[(int64_t)col[p] * n + r] = val[p];
  }
}

int rocblas_check(rocblas_status s, const char* what) {
  if (s == rocblas_status_success) return EDLGT_OK;
  return set_error(s == rocblas_status_memory_error ? EDLGT_ERR_OUT_OF_MEMORY : EDLGT_ERR_HIP,
                   std::string(what) + ": " + rocblas_status_to_string(s));
}

struct Handle {
  rocblas_handle h = nullptr;
  ~Handle() {
    if (h) (void)rocblas_destroy_handle(h);
  }
};

}  // namespace

int csr_dense_eigh(const Csr& h, double* evals, void* evecs, hipStream_t st) {
  const int64_t n = h.n_rows;
  if (h.mf) return set_error(EDLGT_ERR_INVALID_ARG, "not available for a matrix-free operator (no stored entries)");
  if (!h.val.p && h.nnz) return set_error(EDLGT_ERR_INVALID_ARG, "the value array was released (coded matrix)");
  if (h.n_terms != 1) return set_error(EDLGT_ERR_INVALID_ARG, "a multi-term matrix has no single dense form");
  if (n < 1 || n > kDenseEighMaxN)
    return set_error(EDLGT_ERR_INVALID_ARG, "dense eigensolve of " + std::to_string(n) + " states (1.." +
                                                std::to_string(kDenseEighMaxN) + " supported)");
  if (!evals || !evecs) return set_error(EDLGT_ERR_INVALID_ARG, "null output");
  const bool cplx = h.value_type == EDLGT_VALUE_C128;
  if (!cplx && h.value_type != EDLGT_VALUE_F64)
    return set_error(EDLGT_ERR_INVALID_ARG, "dense eigensolve needs complex128 or float64 values");
  const size_t es = cplx ? 16 : 8;
  EDLGT_HIP(hipMemsetAsync(evecs, 0, (size_t)n * n * es, st));
  const unsigned g = (unsigned)std::min<int64_t>((n + kTB - 1) / kTB, 4096);
  if (cplx)
    hipLaunchKernelGGL(csr_to_dense_colmajor<double2>, dim3(g), dim3(kTB), 0, st, h.row_ptr.as<int64_t>(),
                       h.col.as<int32_t>(), h.val.as<double2>(), n, static_cast<double2*>(evecs));
  else
    hipLaunchKernelGGL(csr_to_dense_colmajor<double>, dim3(g), dim3(kTB), 0, st, h.row_ptr.as<int64_t>(),
                       h.col.as<int32_t>(), h.val.as<double>(), n, static_cast<double*>(evecs));
  EDLGT_HIP(hipGetLastError());
  DevBuf d, e, info;
  EDLGT_TRY(d.alloc(n * sizeof(double)));
  EDLGT_TRY(e.alloc(n * sizeof(double)));
  EDLGT_TRY(info.alloc(sizeof(rocblas_int)));
  Handle hd;
  EDLGT_TRY(rocblas_check(rocblas_create_handle(&hd.h), "rocblas_create_handle"));
  EDLGT_TRY(rocblas_check(rocblas_set_stream(hd.h, st), "rocblas_set_stream"));
  const rocblas_int ni = (rocblas_int)n;
  // lower triangle of A (the matrix is Hermitian); eigenvector k overwrites
  // column k, i.e. row k of the row-major (N, N) output
  if (cplx)
    EDLGT_TRY(rocblas_check(rocsolver_zheevd(hd.h, rocblas_evect_original, rocblas_fill_lower, ni,
                                             static_cast<rocblas_double_complex*>(evecs), ni, d.as<double>(),
                                             e.as<double>(), info.as<rocblas_int>()),
                            "rocsolver_zheevd"));
  else
    EDLGT_TRY(rocblas_check(rocsolver_dsyevd(hd.h, rocblas_evect_original, rocblas_fill_lower, ni,
                                             static_cast<double*>(evecs), ni, d.as<double>(), e.as<double>(),
                                             info.as<rocblas_int>()),
                            "rocsolver_dsyevd"));
  rocblas_int inf = 0;
  EDLGT_HIP(hipMemcpyAsync(evals, d.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipMemcpyAsync(&inf, info.p, sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  if (inf != 0)
    return set_error(EDLGT_ERR_NOT_CONVERGED, "dense eigensolve did not converge (info " + std::to_string(inf) + ")");
  return EDLGT_OK;
}

}  // namespace edlgt
