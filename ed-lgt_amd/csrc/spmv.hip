// CSR sparse products y = H x and Y = H X on the device.
//
// Restates csr_spmv / csr_matmat (edlgt/modeling/qmb_hamiltonian.py:705-764):
// y_i = sum_p val[p] * x[col[p]] over the row's stored entries, in storage
// order.  The products are HBM-bound (~0.4 flop/byte in FP64 complex), so the
// kernel is built for bandwidth:
//   * a group of G lanes (G = 4..64, chosen from the mean row length) owns a
//     row, reads the row's col/val with unit-stride (coalesced) loads and
//     gathers x; lane partial sums are combined with a fixed xor-shuffle
//     tree, so results are bitwise reproducible;
//   * matrix entries are streamed with non-temporal loads so that the
//     gathered vector keeps the L2 / MALL capacity;
//   * SpMM reads each entry once for all NB right-hand sides.
// Value/vector types: FP64 real or complex values with FP64 vectors; FP32
// values (mixed precision) with FP64 or FP32 gathered vectors; accumulation
// is always FP64.
#include <algorithm>

#include "common.h"
#include "kernels.h"

namespace edlgt {
namespace {

template <class T> __device__ inline T ldnt(const T* p) { return __builtin_nontemporal_load(p); }

// value loads (non-temporal), widened to FP64
__device__ inline double2 load_val(const double2* v, int64_t p) {
  const double* q = reinterpret_cast<const double*>(v) + 2 * p;
  return make_double2(ldnt(q), ldnt(q + 1));
}
__device__ inline double2 load_val(const float2* v, int64_t p) {
  const float* q = reinterpret_cast<const float*>(v) + 2 * p;
  return make_double2((double)ldnt(q), (double)ldnt(q + 1));
}
__device__ inline double load_val(const double* v, int64_t p) { return ldnt(v + p); }
__device__ inline double load_val(const float* v, int64_t p) { return (double)ldnt(v + p); }

// gathered vector loads, widened to FP64
__device__ inline double2 load_x(const double2* x, int64_t i) { return x[i]; }
__device__ inline double2 load_x(const float2* x, int64_t i) {
  const float2 f = x[i];
  return make_double2((double)f.x, (double)f.y);
}
__device__ inline double load_x(const double* x, int64_t i) { return x[i]; }
__device__ inline double load_x(const float* x, int64_t i) { return (double)x[i]; }

__device__ inline void fma_acc(double2& acc, double2 a, double2 b) {
  acc.x += a.x * b.x - a.y * b.y;
  acc.y += a.x * b.y + a.y * b.x;
}
__device__ inline void fma_acc(double2& acc, double a, double2 b) {
  acc.x += a * b.x;
  acc.y += a * b.y;
}
__device__ inline void fma_acc(double& acc, double a, double b) { acc += a * b; }

template <class A> __device__ inline A zero_acc();
template <> __device__ inline double zero_acc<double>() { return 0.0; }
template <> __device__ inline double2 zero_acc<double2>() { return make_double2(0.0, 0.0); }

template <int G>
__device__ inline double group_sum(double v) {
#pragma unroll
  for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <int G>
__device__ inline double2 group_sum(double2 v) {
  v.x = group_sum<G>(v.x);
  v.y = group_sum<G>(v.y);
  return v;
}

// VT: stored value type; XT: gathered vector element type; AT: accumulator /
// output element type (double or double2); NB right-hand sides, column-major
// with leading dimension ld.
//
// Row partition: the grid is a multiple of the 8 XCDs and workgroup b runs on
// XCD b % 8, so logical block L = (b % 8) * (G / 8) + b / 8 gives every XCD one
// contiguous slab of rows, and each block a contiguous chunk inside it.  Rows
// that are close share most gathered x lines (a term moves a row to nearby
// columns for nearby rows), so those lines are fetched once into the XCD's L2
// instead of once per XCD.
template <class VT, class XT, class AT, int G, int NB>
__global__ void __launch_bounds__(256) csr_spmm_kernel(int64_t n, const int64_t* __restrict__ row_ptr,
                                                       const int32_t* __restrict__ col,
                                                       const VT* __restrict__ val,
                                                       const XT* __restrict__ x, int64_t ldx,
                                                       AT* __restrict__ y, int64_t ldy,
                                                       int64_t rows_per_block) {
  const int sub = threadIdx.x & (G - 1);
  const unsigned nb = gridDim.x;
  const unsigned L = (nb % 8u == 0u) ? (blockIdx.x % 8u) * (nb / 8u) + blockIdx.x / 8u : blockIdx.x;
  const int64_t r_beg = (int64_t)L * rows_per_block;
  const int64_t r_end = min(n, r_beg + rows_per_block);
  for (int64_t r = r_beg + (threadIdx.x / G); r < r_end; r += 256 / G) {
    const int64_t beg = row_ptr[r], end = row_ptr[r + 1];
    AT acc[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = zero_acc<AT>();
    for (int64_t p = beg + sub; p < end; p += G) {
      const int32_t c = ldnt(col + p);
      const auto v = load_val(val, p);
#pragma unroll
      for (int b = 0; b < NB; ++b) fma_acc(acc[b], v, load_x(x + b * ldx, c));
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const AT s = group_sum<G>(acc[b]);
      if (sub == 0) y[b * ldy + r] = s;
    }
  }
}

template <class VT, class XT, class AT, int NB>
int launch_g(int g, int64_t grid, int64_t n, const int64_t* rp, const int32_t* col, const VT* val,
             const XT* x, int64_t ldx, AT* y, int64_t ldy, hipStream_t st) {
  // Default: 8 XCDs x 32 CUs with ~8 resident 256-thread blocks each -> 2048
  // blocks, every block owning a contiguous chunk of rows.
  const int rows_per_pass = 256 / g;
  int64_t blocks = std::min<int64_t>(grid > 0 ? grid : 2048,
                                     std::max<int64_t>(8, (n + rows_per_pass - 1) / rows_per_pass));
  blocks = (blocks + 7) / 8 * 8;
  const int64_t rpb = (n + blocks - 1) / blocks;
#define EDLGT_SPMM_CASE(GG)                                                                        \
  case GG:                                                                                         \
    hipLaunchKernelGGL((csr_spmm_kernel<VT, XT, AT, GG, NB>), dim3((unsigned)blocks), dim3(256), 0, \
                       st, n, rp, col, val, x, ldx, y, ldy, rpb);                                   \
    break;
  switch (g) {
    EDLGT_SPMM_CASE(4)
    EDLGT_SPMM_CASE(8)
    EDLGT_SPMM_CASE(16)
    EDLGT_SPMM_CASE(32)
    EDLGT_SPMM_CASE(64)
    default:
      return set_error(EDLGT_ERR_INVALID_ARG, "unsupported lanes-per-row");
  }
#undef EDLGT_SPMM_CASE
  EDLGT_HIP(hipGetLastError());
  return EDLGT_OK;
}

template <class VT, class XT, class AT>
int launch_nb(int g, int64_t grid, int nb, int64_t n, const int64_t* rp, const int32_t* col,
              const VT* val, const XT* x, int64_t ldx, AT* y, int64_t ldy, hipStream_t st) {
  switch (nb) {
    case 1: return launch_g<VT, XT, AT, 1>(g, grid, n, rp, col, val, x, ldx, y, ldy, st);
    case 2: return launch_g<VT, XT, AT, 2>(g, grid, n, rp, col, val, x, ldx, y, ldy, st);
    case 4: return launch_g<VT, XT, AT, 4>(g, grid, n, rp, col, val, x, ldx, y, ldy, st);
    case 8: return launch_g<VT, XT, AT, 8>(g, grid, n, rp, col, val, x, ldx, y, ldy, st);
  }
  return set_error(EDLGT_ERR_INVALID_ARG, "unsupported block width");
}

}  // namespace

int lanes_per_row(const Csr& h) {
  if (h.n_rows <= 0) return 4;
  const double avg = (double)h.nnz / (double)h.n_rows;
  if (h.g_override) return h.g_override;
  if (avg <= 6) return 4;
  if (avg <= 14) return 8;
  if (avg <= 40) return 16;
  if (avg <= 96) return 32;
  return 64;
}

int csr_spmm(const Csr& h, int precision, int xtype, int n_rhs, const void* x, int64_t ldx, void* y,
             int64_t ldy, hipStream_t st) {
  const int g = lanes_per_row(h);
  const int64_t n = h.n_rows;
  const int64_t* rp = h.row_ptr.as<int64_t>();
  const int32_t* col = h.col.as<int32_t>();
  const bool cplx = h.value_type == EDLGT_VALUE_C128;
  const bool lo = precision == EDLGT_PREC_MIXED;
  if (lo && !h.val_lo.p) return set_error(EDLGT_ERR_INVALID_ARG, "low-precision values not built");
  // Split n_rhs into supported widths (8, 4, 2, 1).
  int done = 0;
  while (done < n_rhs) {
    const int left = n_rhs - done;
    const int nb = left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
    const size_t xe = (xtype == EDLGT_VALUE_C128) ? 16 : (xtype == EDLGT_VALUE_C64) ? 8
                      : (xtype == EDLGT_VALUE_F64) ? 8 : 4;
    const size_t ye = cplx ? 16 : 8;
    const char* xb = static_cast<const char*>(x) + (size_t)done * ldx * xe;
    char* yb = static_cast<char*>(y) + (size_t)done * ldy * ye;
    int s;
    if (cplx) {
      if (!lo && xtype == EDLGT_VALUE_C128)
        s = launch_nb(g, h.grid_override, nb, n, rp, col, h.val.as<double2>(), (const double2*)xb, ldx, (double2*)yb, ldy, st);
      else if (lo && xtype == EDLGT_VALUE_C128)
        s = launch_nb(g, h.grid_override, nb, n, rp, col, h.val_lo.as<float2>(), (const double2*)xb, ldx, (double2*)yb, ldy, st);
      else if (lo && xtype == EDLGT_VALUE_C64)
        s = launch_nb(g, h.grid_override, nb, n, rp, col, h.val_lo.as<float2>(), (const float2*)xb, ldx, (double2*)yb, ldy, st);
      else
        return set_error(EDLGT_ERR_INVALID_ARG, "complex matrix needs complex vectors");
    } else {
      if (!lo && xtype == EDLGT_VALUE_F64)
        s = launch_nb(g, h.grid_override, nb, n, rp, col, h.val.as<double>(), (const double*)xb, ldx, (double*)yb, ldy, st);
      else if (!lo && xtype == EDLGT_VALUE_C128)
        s = launch_nb(g, h.grid_override, nb, n, rp, col, h.val.as<double>(), (const double2*)xb, ldx, (double2*)yb, ldy, st);
      else if (lo && xtype == EDLGT_VALUE_F64)
        s = launch_nb(g, h.grid_override, nb, n, rp, col, h.val_lo.as<float>(), (const double*)xb, ldx, (double*)yb, ldy, st);
      else if (lo && xtype == EDLGT_VALUE_F32)
        s = launch_nb(g, h.grid_override, nb, n, rp, col, h.val_lo.as<float>(), (const float*)xb, ldx, (double*)yb, ldy, st);
      else if (lo && xtype == EDLGT_VALUE_C128)
        s = launch_nb(g, h.grid_override, nb, n, rp, col, h.val_lo.as<float>(), (const double2*)xb, ldx, (double2*)yb, ldy, st);
      else
        return set_error(EDLGT_ERR_INVALID_ARG, "unsupported vector type for a real matrix");
    }
    if (s != EDLGT_OK) return s;
    done += nb;
  }
  return EDLGT_OK;
}

namespace {
__global__ void to_f32(const double* __restrict__ in, float* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}
}  // namespace

int convert_f64_to_f32(const double* in, float* out, int64_t n, hipStream_t st) {
  if (n <= 0) return EDLGT_OK;
  hipLaunchKernelGGL(to_f32, dim3(grid_for(n, 256)), dim3(256), 0, st, in, out, n);
  EDLGT_HIP(hipGetLastError());
  return EDLGT_OK;
}

int csr_make_low_precision(Csr& h, hipStream_t st) {
  if (h.val_lo.p) return EDLGT_OK;
  const int64_t ncomp = h.nnz * (h.value_type == EDLGT_VALUE_C128 ? 2 : 1);
  EDLGT_TRY(h.val_lo.alloc(std::max<int64_t>(ncomp, 1) * sizeof(float)));
  EDLGT_TRY(convert_f64_to_f32(h.val.as<double>(), h.val_lo.as<float>(), ncomp, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  return EDLGT_OK;
}

int csr_autotune(Csr& h, int precision, int xtype, int n_rhs, const void* x, void* y,
                 hipStream_t st, int32_t* best_g, int64_t* best_grid, float* best_ms) {
  const double avg = h.n_rows > 0 ? (double)h.nnz / (double)h.n_rows : 1.0;
  const int gs[] = {4, 8, 16, 32, 64};
  const int64_t grids[] = {1024, 2048, 4096, 8192, 16384};
  hipEvent_t e0, e1;
  EDLGT_HIP(hipEventCreate(&e0));
  EDLGT_HIP(hipEventCreate(&e1));
  const int32_t save_g = h.g_override;
  const int64_t save_grid = h.grid_override;
  float best = 1e30f;
  int32_t bg = 0;
  int64_t bgrid = 0;
  int status = EDLGT_OK;
  for (int g : gs) {
    if (g > 8 && g > 2 * avg) continue;
    for (int64_t grid : grids) {
      h.g_override = g;
      h.grid_override = grid;
      status = csr_spmm(h, precision, xtype, n_rhs, x, h.n_rows, y, h.n_rows, st);  // warm
      if (status != EDLGT_OK) break;
      (void)hipEventRecord(e0, st);
      const int reps = 3;
      for (int r = 0; r < reps && status == EDLGT_OK; ++r)
        status = csr_spmm(h, precision, xtype, n_rhs, x, h.n_rows, y, h.n_rows, st);
      (void)hipEventRecord(e1, st);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      ms /= reps;
      if (status == EDLGT_OK && ms < best) {
        best = ms;
        bg = g;
        bgrid = grid;
      }
    }
    if (status != EDLGT_OK) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (status != EDLGT_OK) {
    h.g_override = save_g;
    h.grid_override = save_grid;
    return status;
  }
  h.g_override = bg;
  h.grid_override = bgrid;
  if (best_g) *best_g = bg;
  if (best_grid) *best_grid = bgrid;
  if (best_ms) *best_ms = best;
  return EDLGT_OK;
}

}  // namespace edlgt
