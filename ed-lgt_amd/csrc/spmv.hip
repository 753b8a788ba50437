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

__device__ inline double2 operator+(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

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

// Vector-CSR product (G lanes per row), used for several right-hand sides.
// VT: stored value type; XT: gathered vector element type; AT: accumulator /
// output element type (double or double2); NB right-hand sides, column-major
// with leading dimension ld.  Blocks take row groups in grid-stride order so
// that the blocks in flight cover one compact window of rows: the gathered x
// lines of neighbouring rows are then reused from L2 / MALL (measured faster
// on MI355X than giving each XCD its own contiguous slab of rows).
template <class VT, class XT, class AT, int G, int NB>
__global__ void __launch_bounds__(256) csr_spmm_kernel(int64_t n, const int64_t* __restrict__ row_ptr,
                                                       const int32_t* __restrict__ col,
                                                       const VT* __restrict__ val,
                                                       const XT* __restrict__ x, int64_t ldx,
                                                       AT* __restrict__ y, int64_t ldy) {
  const int sub = threadIdx.x & (G - 1);
  const int64_t stride = (int64_t)gridDim.x * (256 / G);
  for (int64_t r = (int64_t)blockIdx.x * (256 / G) + (threadIdx.x / G); r < n; r += stride) {
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

// CSR-stream product for short rows (one right-hand side).
//
// The nonzeros are cut into tiles of TILE entries; tile k's block owns the
// rows that start inside [k*TILE, (k+1)*TILE) (row_blocks[k] =
// lower_bound(row_ptr, k*TILE), found on the device).  A BS-thread block
// streams its contiguous entries with coalesced non-temporal loads, each
// thread keeping U independent gathers of x in flight, writes the products
// to LDS, and then reduces every row sequentially from LDS (fixed order ->
// bitwise reproducible).  Blocks run in row order, so the whole chip sweeps a
// compact window of rows and the gathered x lines are reused from L2 / MALL.
// A block whose entries exceed the LDS capacity (a very long row) falls back
// to a block-cooperative reduction per row.  (BS, TILE, U) is a launch
// variant chosen per matrix by the autotuner; (512, 2048, 4) is the default
// and was the fastest on every matrix measured on MI355X.
constexpr int kBS = 512;     // default threads per block of the stream kernel
constexpr int kTile = 2048;  // default nonzeros per tile

// 16-byte non-temporal value loads (one dwordx4 instead of two dwordx2)
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));
__device__ inline double2 load_val16(const double2* v, int64_t p) {
  const dvec2 t = __builtin_nontemporal_load(reinterpret_cast<const dvec2*>(v) + p);
  return make_double2(t.x, t.y);
}
__device__ inline double2 load_val16(const float2* v, int64_t p) {
  const fvec2 t = __builtin_nontemporal_load(reinterpret_cast<const fvec2*>(v) + p);
  return make_double2((double)t.x, (double)t.y);
}
__device__ inline double load_val16(const double* v, int64_t p) { return ldnt(v + p); }
__device__ inline double load_val16(const float* v, int64_t p) { return (double)ldnt(v + p); }

template <int G>
__device__ inline double sum_lanes(double v) {
#pragma unroll
  for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <int G>
__device__ inline double2 sum_lanes(double2 v) {
  return make_double2(sum_lanes<G>(v.x), sum_lanes<G>(v.y));
}

// Row sums from the LDS products with G lanes per row (fixed per block).
template <int G, class AT, int BS = kBS>
__device__ inline void reduce_rows(const AT* prod, const int64_t* __restrict__ row_ptr, int64_t r0,
                                   int64_t nr, int64_t p0, AT* __restrict__ y) {
  const int sub = threadIdx.x & (G - 1);
  constexpr int per = BS / G;
  for (int64_t i = threadIdx.x / G; i < (nr + per - 1) / per * per; i += per) {
    AT s = zero_acc<AT>();
    if (i < nr) {
      const int64_t b = row_ptr[r0 + i] - p0, e = row_ptr[r0 + i + 1] - p0;
      for (int64_t q = b + sub; q < e; q += G) s = s + prod[q];
    }
    s = sum_lanes<G>(s);
    if (sub == 0 && i < nr) y[r0 + i] = s;
  }
}

// Tile of a stream-kernel block.  The hardware hands workgroup b to XCD
// b % 8 (8 XCDs, each with its own L2); this gives every XCD runs of
// kXcdRun consecutive tiles instead of every eighth tile, so rows that share
// near-diagonal x lines run on one L2 (measured on QED 4x3: 3.13 -> 3.04 ms;
// runs of 4 and 16 were slower than 8).
constexpr int64_t kXcdRun = 8;
__device__ inline int64_t stream_tile(int64_t b, int64_t n_tiles) {
  constexpr int64_t per = 8 * kXcdRun;
  const int64_t full = n_tiles / per * per;
  if (b >= full) return b;
  const int64_t r = b % per;
  return b / per * per + (r % 8) * kXcdRun + r / 8;
}

template <class VT, class XT, class AT, int BS, int TILE, int U>
__global__ void __launch_bounds__(BS) csr_stream_kernel(
    int64_t n, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const VT* __restrict__ val, const XT* __restrict__ x, AT* __restrict__ y,
    const int64_t* __restrict__ row_blocks, const int64_t* __restrict__ block_pos) {
  constexpr int CAP = TILE + 256;
  __shared__ AT prod[CAP];
  const int64_t t = stream_tile(blockIdx.x, gridDim.x);
  const int64_t r0 = row_blocks[t], r1 = row_blocks[t + 1];
  if (r0 >= r1) return;
  const int64_t p0 = block_pos[t], p1 = block_pos[t + 1];
  const int64_t cnt = p1 - p0;
  const int64_t nr = r1 - r0;
  const int tid = threadIdx.x;
  if (cnt <= CAP) {
    // memory phase at raised wave priority, so that the waves of blocks still
    // loading issue ahead of those reducing from LDS (measured: -0.3 %)
    __builtin_amdgcn_s_setprio(2);
    for (int64_t base = 0; base < cnt; base += BS * U) {
      int32_t c[U];
      decltype(load_val16(val, 0)) v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = base + u * BS + tid;
        if (i < cnt) {
          c[u] = ldnt(col + p0 + i);
          v[u] = load_val16(val, p0 + i);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = base + u * BS + tid;
        if (i < cnt) {
          AT a = zero_acc<AT>();
          fma_acc(a, v[u], load_x(x, c[u]));
          prod[i] = a;
        }
      }
    }
    __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    if (nr <= BS / 8) reduce_rows<8, AT, BS>(prod, row_ptr, r0, nr, p0, y);
    else if (nr <= BS / 4) reduce_rows<4, AT, BS>(prod, row_ptr, r0, nr, p0, y);
    else if (nr <= BS / 2) reduce_rows<2, AT, BS>(prod, row_ptr, r0, nr, p0, y);
    else reduce_rows<1, AT, BS>(prod, row_ptr, r0, nr, p0, y);
  } else {
    // long rows: all threads on one row at a time, fixed-shape tree sum
    AT* red = prod;  // CAP >= BS
    for (int64_t r = r0; r < r1; ++r) {
      AT s = zero_acc<AT>();
      for (int64_t p = row_ptr[r] + tid; p < row_ptr[r + 1]; p += BS)
        fma_acc(s, load_val16(val, p), load_x(x, ldnt(col + p)));
      red[tid] = s;
      __syncthreads();
      for (int o = BS / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] = red[tid] + red[tid + o];
        __syncthreads();
      }
      if (tid == 0) y[r] = red[0];
      __syncthreads();
    }
  }
}

__global__ void stream_block_pos(const int64_t* __restrict__ row_ptr, const int64_t* __restrict__ rb,
                                 int64_t n_blocks, int64_t* __restrict__ pos) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= n_blocks;
       k += (int64_t)gridDim.x * blockDim.x)
    pos[k] = row_ptr[rb[k]];
}

__global__ void stream_row_blocks(const int64_t* __restrict__ row_ptr, int64_t n, int64_t n_blocks,
                                  int64_t* __restrict__ rb, int64_t tile) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= n_blocks;
       k += (int64_t)gridDim.x * blockDim.x) {
    if (k == n_blocks) {
      rb[k] = n;
      continue;
    }
    const int64_t target = k * tile;
    int64_t lo = 0, len = n;  // first row with row_ptr[r] >= target
    while (len > 0) {
      const int64_t half = len >> 1;
      if (row_ptr[lo + half] < target) {
        lo += half + 1;
        len -= half + 1;
      } else {
        len = half;
      }
    }
    rb[k] = lo;
  }
}

// ---- dictionary-coded values (csr_compress): entry p carries code[p] and
// its value is tab[code[p]], a table already multiplied by the product's term
// coefficients.  Same loop structure and summation order as the kernels
// above; the streamed bytes per entry drop to 4 (col) + 1 or 2 (code).
constexpr int kTabLds = 1024;      // tables up to this many entries sit in LDS
constexpr int kRowsTabLds = 2048;  // NB x n_tab bound for the multi-RHS kernel


template <class TT, class CT, class XT, class AT, bool kLds, int BS, int TILE, int U>
__global__ void __launch_bounds__(BS) csr_stream_coded(
    int64_t n, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const CT* __restrict__ code, const TT* __restrict__ tab, int n_tab, const XT* __restrict__ x,
    AT* __restrict__ y, const int64_t* __restrict__ row_blocks, const int64_t* __restrict__ block_pos) {
  constexpr int CAP = TILE + 256;
  __shared__ AT prod[CAP];
  __shared__ TT stab[kLds ? kTabLds : 1];
  if (kLds) {
    for (int i = threadIdx.x; i < n_tab; i += BS) stab[i] = tab[i];
    __syncthreads();
  }
  const TT* tb = kLds ? stab : tab;
  const int64_t t = stream_tile(blockIdx.x, gridDim.x);
  const int64_t r0 = row_blocks[t], r1 = row_blocks[t + 1];
  if (r0 >= r1) return;
  const int64_t p0 = block_pos[t], p1 = block_pos[t + 1];
  const int64_t cnt = p1 - p0;
  const int64_t nr = r1 - r0;
  const int tid = threadIdx.x;
  if (cnt <= CAP) {
    __builtin_amdgcn_s_setprio(2);
    for (int64_t base = 0; base < cnt; base += BS * U) {
      int32_t c[U];
      int k[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = base + u * BS + tid;
        if (i < cnt) {
          c[u] = ldnt(col + p0 + i);
          k[u] = (int)ldnt(code + p0 + i);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = base + u * BS + tid;
        if (i < cnt) {
          AT a = zero_acc<AT>();
          fma_acc(a, tb[k[u]], load_x(x, c[u]));
          prod[i] = a;
        }
      }
    }
    __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    if (nr <= BS / 8) reduce_rows<8, AT, BS>(prod, row_ptr, r0, nr, p0, y);
    else if (nr <= BS / 4) reduce_rows<4, AT, BS>(prod, row_ptr, r0, nr, p0, y);
    else if (nr <= BS / 2) reduce_rows<2, AT, BS>(prod, row_ptr, r0, nr, p0, y);
    else reduce_rows<1, AT, BS>(prod, row_ptr, r0, nr, p0, y);
  } else {
    AT* red = prod;
    for (int64_t r = r0; r < r1; ++r) {
      AT s = zero_acc<AT>();
      for (int64_t p = row_ptr[r] + tid; p < row_ptr[r + 1]; p += BS)
        fma_acc(s, tb[(int)ldnt(code + p)], load_x(x, ldnt(col + p)));
      red[tid] = s;
      __syncthreads();
      for (int o = BS / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] = red[tid] + red[tid + o];
        __syncthreads();
      }
      if (tid == 0) y[r] = red[0];
      __syncthreads();
    }
  }
}

template <class VT, class XT, class AT, int NB>
int launch_g(int g, int64_t grid, int64_t n, const int64_t* rp, const int32_t* col, const VT* val,
             const XT* x, int64_t ldx, AT* y, int64_t ldy, hipStream_t st) {
  const int rows_per_pass = 256 / g;
  const int64_t blocks = std::min<int64_t>(grid > 0 ? grid : 65536,
                                           std::max<int64_t>(1, (n + rows_per_pass - 1) / rows_per_pass));
#define EDLGT_SPMM_CASE(GG)                                                                        \
  case GG:                                                                                         \
    hipLaunchKernelGGL((csr_spmm_kernel<VT, XT, AT, GG, NB>), dim3((unsigned)blocks), dim3(256), 0, \
                       st, n, rp, col, val, x, ldx, y, ldy);                                        \
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

int tile_index(int tile);

// Launch variants of both stream kernels (threads per block, nonzeros per
// tile, independent gathers in flight per thread), picked per matrix by
// csr_autotune (plain_variant / coded_variant).
struct StreamVariant {
  int bs, tile, unroll;
};
constexpr StreamVariant kStreamVariants[] = {{512, 2048, 4}, {256, 1024, 4}, {512, 4096, 8}, {1024, 4096, 4},
                                             {256, 2048, 8}};
constexpr int kNumStreamVariants = 5;

template <class VT, class XT, class AT>
int launch_stream(const Csr& h, const VT* val, const XT* x, AT* y, hipStream_t st) {
  const int v = h.plain_variant;
  const int ti = tile_index(kStreamVariants[v].tile);
  if (!h.tile_blocks[ti].p) return set_error(EDLGT_ERR_INVALID_ARG, "stream tiles not prepared");
  const unsigned nb = (unsigned)h.n_tile_blocks[ti];
  const int64_t* rb = h.tile_blocks[ti].as<int64_t>();
  const int64_t* pos = h.tile_pos[ti].as<int64_t>();
#define EDLGT_SP(BS, TILE, U)                                                                     \
  hipLaunchKernelGGL((csr_stream_kernel<VT, XT, AT, BS, TILE, U>), dim3(nb), dim3(BS), 0, st, h.n_rows, \
                     h.row_ptr.as<int64_t>(), h.col.as<int32_t>(), val, x, y, rb, pos)
  switch (v) {
    case 0: EDLGT_SP(512, 2048, 4); break;
    case 1: EDLGT_SP(256, 1024, 4); break;
    case 2: EDLGT_SP(512, 4096, 8); break;
    case 3: EDLGT_SP(1024, 4096, 4); break;
    default: EDLGT_SP(256, 2048, 8); break;
  }
#undef EDLGT_SP
  EDLGT_HIP(hipGetLastError());
  return EDLGT_OK;
}


int spmv_coded(const Csr& h, const void* tab, int xtype, const void* xb, void* yb, hipStream_t st);
int tile_index(int tile) { return tile == 1024 ? 0 : tile == 2048 ? 1 : 2; }

// Row-start partition of the nonzeros into tiles of `tile` entries (the
// stream kernels' work units): tile_blocks[k] = first row starting at or
// after k * tile, tile_pos[k] = its first entry.
int csr_prepare_tiles(Csr& h, int tile, hipStream_t st) {
  const int ti = tile_index(tile);
  if (h.tile_blocks[ti].p) return EDLGT_OK;
  const int64_t nb = std::max<int64_t>(1, (h.nnz + tile - 1) / tile);
  if (nb > 0x7fffffffLL) return set_error(EDLGT_ERR_INDEX_OVERFLOW, "too many stream blocks");
  EDLGT_TRY(h.tile_blocks[ti].alloc((nb + 1) * sizeof(int64_t)));
  hipLaunchKernelGGL(stream_row_blocks, dim3(grid_for(nb + 1, 256)), dim3(256), 0, st,
                     h.row_ptr.as<int64_t>(), h.n_rows, nb, h.tile_blocks[ti].as<int64_t>(), (int64_t)tile);
  EDLGT_HIP(hipGetLastError());
  EDLGT_TRY(h.tile_pos[ti].alloc((nb + 1) * sizeof(int64_t)));
  hipLaunchKernelGGL(stream_block_pos, dim3(grid_for(nb + 1, 256)), dim3(256), 0, st,
                     h.row_ptr.as<int64_t>(), h.tile_blocks[ti].as<int64_t>(), nb, h.tile_pos[ti].as<int64_t>());
  EDLGT_HIP(hipGetLastError());
  EDLGT_HIP(hipStreamSynchronize(st));
  h.n_tile_blocks[ti] = nb;
  return EDLGT_OK;
}

int csr_prepare_stream(Csr& h, hipStream_t st) {
  EDLGT_TRY(csr_prepare_tiles(h, kTile, st));
  h.n_stream_blocks = h.n_tile_blocks[tile_index(kTile)];
  return EDLGT_OK;
}

// Times the coded stream variants on x -> y (one right-hand side, FP64
// tables) and keeps the fastest in h.coded_variant.
int coded_autotune(Csr& h, int xtype, const void* x, void* y, hipStream_t st, float* best_ms) {
  hipEvent_t e0, e1;
  EDLGT_HIP(hipEventCreate(&e0));
  EDLGT_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int bv = 0, status = EDLGT_OK;
  for (int v = 0; v < kNumStreamVariants && status == EDLGT_OK; ++v) {
    status = csr_prepare_tiles(h, kStreamVariants[v].tile, st);
    if (status != EDLGT_OK) break;
    h.coded_variant = v;
    status = spmv_coded(h, h.tab_unit.p, xtype, x, y, st);  // warm
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < 3 && status == EDLGT_OK; ++r) status = spmv_coded(h, h.tab_unit.p, xtype, x, y, st);
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms / 3 < best) {
      best = ms / 3;
      bv = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  h.coded_variant = bv;
  if (best_ms) *best_ms = best;
  return status;
}

// Single right-hand side through the CSR-stream kernel.
int spmv_stream(const Csr& h, int precision, int xtype, const void* xb, void* yb, hipStream_t st) {
  const bool cplx = h.value_type == EDLGT_VALUE_C128;
  const bool lo = precision == EDLGT_PREC_MIXED;
  if (cplx) {
    if (!lo && xtype == EDLGT_VALUE_C128)
      return launch_stream(h, h.val.as<double2>(), (const double2*)xb, (double2*)yb, st);
    if (lo && xtype == EDLGT_VALUE_C128)
      return launch_stream(h, h.val_lo.as<float2>(), (const double2*)xb, (double2*)yb, st);
    if (lo && xtype == EDLGT_VALUE_C64)
      return launch_stream(h, h.val_lo.as<float2>(), (const float2*)xb, (double2*)yb, st);
    return set_error(EDLGT_ERR_INVALID_ARG, "complex matrix needs complex vectors");
  }
  if (!lo && xtype == EDLGT_VALUE_F64)
    return launch_stream(h, h.val.as<double>(), (const double*)xb, (double*)yb, st);
  if (!lo && xtype == EDLGT_VALUE_C128)
    return launch_stream(h, h.val.as<double>(), (const double2*)xb, (double2*)yb, st);
  if (lo && xtype == EDLGT_VALUE_F64)
    return launch_stream(h, h.val_lo.as<float>(), (const double*)xb, (double*)yb, st);
  if (lo && xtype == EDLGT_VALUE_F32)
    return launch_stream(h, h.val_lo.as<float>(), (const float*)xb, (double*)yb, st);
  if (lo && xtype == EDLGT_VALUE_C128)
    return launch_stream(h, h.val_lo.as<float>(), (const double2*)xb, (double2*)yb, st);
  return set_error(EDLGT_ERR_INVALID_ARG, "unsupported vector type for a real matrix");
}


int tile_index(int tile);

template <class TT, class XT, class AT>
int launch_stream_coded(const Csr& h, const TT* tab, const XT* x, AT* y, hipStream_t st) {
  const bool lds = h.n_dict <= kTabLds;
  const int v = h.coded_variant;
  const int ti = tile_index(kStreamVariants[v].tile);
  if (!h.tile_blocks[ti].p) return set_error(EDLGT_ERR_INVALID_ARG, "stream tiles not prepared");
  const unsigned nb = (unsigned)h.n_tile_blocks[ti];
  const int64_t* rb = h.tile_blocks[ti].as<int64_t>();
  const int64_t* pos = h.tile_pos[ti].as<int64_t>();
#define EDLGT_SC(CT, LDS, BS, TILE, U)                                                              \
  hipLaunchKernelGGL((csr_stream_coded<TT, CT, XT, AT, LDS, BS, TILE, U>), dim3(nb), dim3(BS), 0, st, \
                     h.n_rows, h.row_ptr.as<int64_t>(), h.col.as<int32_t>(), h.code.as<CT>(), tab,   \
                     h.n_dict, x, y, rb, pos)
#define EDLGT_SCV(CT, LDS)                                          \
  switch (v) {                                                      \
    case 0: EDLGT_SC(CT, LDS, 512, 2048, 4); break;                 \
    case 1: EDLGT_SC(CT, LDS, 256, 1024, 4); break;                 \
    case 2: EDLGT_SC(CT, LDS, 512, 4096, 8); break;                 \
    case 3: EDLGT_SC(CT, LDS, 1024, 4096, 4); break;                \
    default: EDLGT_SC(CT, LDS, 256, 2048, 8); break;                \
  }
  if (h.code_bytes == 1) {
    if (lds) { EDLGT_SCV(uint8_t, true) } else { EDLGT_SCV(uint8_t, false) }
  } else {
    if (lds) { EDLGT_SCV(uint16_t, true) } else { EDLGT_SCV(uint16_t, false) }
  }
#undef EDLGT_SCV
#undef EDLGT_SC
  EDLGT_HIP(hipGetLastError());
  return EDLGT_OK;
}

// One right-hand side of a coded matrix; tab: scaled table (double for a real
// matrix, double2 for a complex one).
int spmv_coded(const Csr& h, const void* tab, int xtype, const void* xb, void* yb, hipStream_t st) {
  if (h.value_type == EDLGT_VALUE_C128) {
    if (xtype == EDLGT_VALUE_C128)
      return launch_stream_coded(h, (const double2*)tab, (const double2*)xb, (double2*)yb, st);
    if (xtype == EDLGT_VALUE_C64)  // mixed precision: complex64 gathered vector, FP64 values and sums
      return launch_stream_coded(h, (const double2*)tab, (const float2*)xb, (double2*)yb, st);
    return set_error(EDLGT_ERR_INVALID_ARG, "complex matrix needs complex vectors");
  }
  if (xtype == EDLGT_VALUE_F64)
    return launch_stream_coded(h, (const double*)tab, (const double*)xb, (double*)yb, st);
  if (xtype == EDLGT_VALUE_C128)
    return launch_stream_coded(h, (const double*)tab, (const double2*)xb, (double2*)yb, st);
  if (xtype == EDLGT_VALUE_F32)
    return launch_stream_coded(h, (const double*)tab, (const float*)xb, (double*)yb, st);
  return set_error(EDLGT_ERR_INVALID_ARG, "unsupported vector type for a coded matrix");
}

int spmm_packed(const Csr& h, int xtype, int n_rhs, const void* x, int64_t ldx, void* y, int64_t ldy,
                void* pack, hipStream_t st);

int csr_spmm(const Csr& h, int precision, int xtype, int n_rhs, const void* x, int64_t ldx, void* y,
             int64_t ldy, hipStream_t st, void* pack, size_t pack_bytes) {
  if (h.mf) {
    if (precision != EDLGT_PREC_FP64) return set_error(EDLGT_ERR_INVALID_ARG, "matrix-free products are FP64");
    return matfree_spmm(h, xtype, n_rhs, x, ldx, y, ldy, st);
  }
  if (n_rhs > 1 && precision == EDLGT_PREC_FP64 && pack && pack_bytes >= spmm_pack_bytes(h, xtype) &&
      spmm_packable(h, xtype))
    return spmm_packed(h, xtype, n_rhs, x, ldx, y, ldy, pack, st);
  if (h.n_dict > 0 && h.stream_blocks.p) {
    // coded values: one stream pass per right-hand side (all term coefficients
    // 1); with a reduced-precision vector only the gathers are FP32
    const size_t xe = xtype == EDLGT_VALUE_C128 ? 16 : xtype == EDLGT_VALUE_F32 ? 4 : 8;
    const size_t ye = h.value_type == EDLGT_VALUE_C128 || xtype == EDLGT_VALUE_C128 ? 16 : 8;
    for (int j = 0; j < n_rhs; ++j)
      EDLGT_TRY(spmv_coded(h, h.tab_unit.p, xtype, static_cast<const char*>(x) + (size_t)j * ldx * xe,
                           static_cast<char*>(y) + (size_t)j * ldy * ye, st));
    return EDLGT_OK;
  }
  if (n_rhs == 1 && h.g_override <= 0 && h.stream_blocks.p) {
    if (precision == EDLGT_PREC_MIXED && !h.val_lo.p)
      return set_error(EDLGT_ERR_INVALID_ARG, "low-precision values not built");
    return spmv_stream(h, precision, xtype, x, y, st);
  }
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
  if (h.mf) return set_error(EDLGT_ERR_INVALID_ARG, "a matrix-free operator has no stored values (FP64 only)");
  if (h.n_dict > 0) return EDLGT_OK;  // coded values: products read the exact table
  if (h.val_lo.p) return EDLGT_OK;
  const int64_t ncomp = h.nnz * (h.value_type == EDLGT_VALUE_C128 ? 2 : 1);
  EDLGT_TRY(h.val_lo.alloc(std::max<int64_t>(ncomp, 1) * sizeof(float)));
  EDLGT_TRY(convert_f64_to_f32(h.val.as<double>(), h.val_lo.as<float>(), ncomp, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  return EDLGT_OK;
}

namespace {

// Multi-coefficient product: for right-hand side b, y_b = sum_k coef[b][k] H_k x_b
// where H = sum_k H_k shares one CSR pattern and every entry carries its term
// id k.  The nb right-hand sides are first packed row-interleaved
// (xi[r * NB + b] = x_b[r]), so that the gather for one stored entry fetches
// the NB values of its column as one contiguous line instead of NB scattered
// words from NB separate vectors; one pass then streams col/val/term once for
// all NB right-hand sides (the fused SpMM of batched parameter sweeps and
// served solves).  Each column's accumulation order is the same for any NB
// and any batch composition, so a right-hand side gets the same bits alone or
// in a batch.
constexpr int kMaxTerms = 16;

__device__ inline double2 scale(double c, double2 v) { return make_double2(c * v.x, c * v.y); }
__device__ inline double scale(double c, double v) { return c * v; }

// Columns of a block of vectors: a device array of per-column pointers, or
// column j at base + j * ld.
template <class S>
struct InCols {
  const S* const* ptr;
  const S* base;
  int64_t ld;
  __device__ const S* col(int j) const { return ptr ? ptr[j] : base + (int64_t)j * ld; }
};
template <class S>
struct OutCols {
  S* const* ptr;
  S* base;
  int64_t ld;
  __device__ S* col(int j) const { return ptr ? ptr[j] : base + (int64_t)j * ld; }
};

template <class S, int NB>
__global__ void __launch_bounds__(256) pack_rows(InCols<S> xs, int nb, int64_t n, S* __restrict__ xi) {
  const S* xp[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) xp[b] = b < nb ? xs.col(b) : nullptr;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int b = 0; b < NB; ++b) xi[r * NB + b] = b < nb ? xp[b][r] : zero_acc<S>();
  }
}

// NB consecutive vector elements of one packed row, 16-byte loads.
template <int NB>
__device__ inline void load_row(const double* __restrict__ q, double (&x)[NB]) {
  if constexpr (NB == 1) {
    x[0] = q[0];
  } else {
    const double2* q2 = reinterpret_cast<const double2*>(q);
#pragma unroll
    for (int i = 0; i < NB / 2; ++i) {
      const double2 v = q2[i];
      x[2 * i] = v.x;
      x[2 * i + 1] = v.y;
    }
  }
}
template <int NB>
__device__ inline void load_row(const double2* __restrict__ q, double2 (&x)[NB]) {
#pragma unroll
  for (int i = 0; i < NB; ++i) x[i] = q[i];
}

template <class VT, class S, int G, int NB>
__global__ void __launch_bounds__(256) csr_spmm_coef_rows(
    int64_t n, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const VT* __restrict__ val, const uint8_t* __restrict__ term, const S* __restrict__ xi,
    OutCols<S> ys, int nb, const double* __restrict__ coef) {
  // coef == nullptr: every coefficient 1 (a plain block product)
  __shared__ double cs[NB * kMaxTerms];
  for (int i = threadIdx.x; i < NB * kMaxTerms; i += blockDim.x) cs[i] = coef ? coef[i] : 1.0;
  __syncthreads();
  const int sub = threadIdx.x & (G - 1);
  const int64_t stride = (int64_t)gridDim.x * (256 / G);
  for (int64_t r = (int64_t)blockIdx.x * (256 / G) + (threadIdx.x / G); r < n; r += stride) {
    const int64_t beg = row_ptr[r], end = row_ptr[r + 1];
    S acc[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = zero_acc<S>();
    for (int64_t p = beg + sub; p < end; p += G) {
      const int32_t c = ldnt(col + p);
      const auto v = load_val(val, p);
      const int t = term ? (int)ldnt(term + p) : 0;
      S x[NB];
      load_row<NB>(xi + (int64_t)c * NB, x);
#pragma unroll
      for (int b = 0; b < NB; ++b) fma_acc(acc[b], scale(cs[b * kMaxTerms + t], v), x[b]);
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const S s = group_sum<G>(acc[b]);
      if (sub == 0 && b < nb) ys.col(b)[r] = s;
    }
  }
}

// Per-right-hand-side tables of a coded matrix: out[b][j] = coef[b][term_j] * value_j.
__device__ inline void set_tab(double& o, double c, double2 v) { o = c * v.x; }
__device__ inline void set_tab(double2& o, double c, double2 v) { o = scale(c, v); }

template <class TT>
__global__ void scale_dict(const double2* __restrict__ dval, const int32_t* __restrict__ dterm, int n_dict,
                           const double* __restrict__ coef, int nb, TT* __restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb * n_dict; i += gridDim.x * blockDim.x) {
    const int b = i / n_dict, j = i % n_dict;
    set_tab(out[i], coef[b * kMaxTerms + dterm[j]], dval[j]);
  }
}

template <class TT, class CT, class S, int G, int NB, bool kLds>
__global__ void __launch_bounds__(256) csr_rows_coded(
    int64_t n, const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const CT* __restrict__ code, const TT* __restrict__ tab, int n_tab, int tab_stride,
    const S* __restrict__ xi, OutCols<S> ys, int nb) {
  // column b's value of code k: tab[b * tab_stride + k] (tab_stride 0: one
  // table shared by every column)
  __shared__ TT stab[kLds ? kRowsTabLds : 1];
  if (kLds) {
    for (int i = threadIdx.x; i < (tab_stride ? nb : 1) * n_tab; i += blockDim.x) stab[i] = tab[i];
    __syncthreads();
  }
  const TT* tb = kLds ? stab : tab;
  const int sub = threadIdx.x & (G - 1);
  const int64_t stride = (int64_t)gridDim.x * (256 / G);
  for (int64_t r = (int64_t)blockIdx.x * (256 / G) + (threadIdx.x / G); r < n; r += stride) {
    const int64_t beg = row_ptr[r], end = row_ptr[r + 1];
    S acc[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = zero_acc<S>();
    for (int64_t p = beg + sub; p < end; p += G) {
      const int32_t c = ldnt(col + p);
      const int k = (int)ldnt(code + p);
      S x[NB];
      load_row<NB>(xi + (int64_t)c * NB, x);
#pragma unroll
      for (int b = 0; b < NB; ++b) fma_acc(acc[b], tb[b * tab_stride + k], x[b]);
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const S s = group_sum<G>(acc[b]);
      if (sub == 0 && b < nb) ys.col(b)[r] = s;
    }
  }
}

template <class TT, class S, int G>
int launch_rows_coded(int nb, int nbp, const Csr& h, const TT* tab, int tab_stride, InCols<S> xs,
                      OutCols<S> ys, void* pack, hipStream_t st) {
  const int64_t n = h.n_rows;
  const unsigned pgrid = grid_for(n, 256, 65536);
  const int64_t blocks = std::min<int64_t>(65536, std::max<int64_t>(1, (n + 256 / G - 1) / (256 / G)));
#define EDLGT_RC(CT, NBC, LDS)                                                                      \
  hipLaunchKernelGGL((csr_rows_coded<TT, CT, S, G, NBC, LDS>), dim3((unsigned)blocks), dim3(256), 0, st, \
                     n, h.row_ptr.as<int64_t>(), h.col.as<int32_t>(), h.code.as<CT>(), tab, h.n_dict,  \
                     tab_stride, (const S*)pack, ys, nb)
#define EDLGT_RC_NB(NBC)                                                                            \
  case NBC:                                                                                         \
    hipLaunchKernelGGL((pack_rows<S, NBC>), dim3(pgrid), dim3(256), 0, st, xs, nb, n, (S*)pack);    \
    if (h.code_bytes == 1) {                                                                        \
      if (lds) EDLGT_RC(uint8_t, NBC, true); else EDLGT_RC(uint8_t, NBC, false);                    \
    } else {                                                                                        \
      if (lds) EDLGT_RC(uint16_t, NBC, true); else EDLGT_RC(uint16_t, NBC, false);                  \
    }                                                                                               \
    break;
  const bool lds = (tab_stride ? nbp : 1) * h.n_dict <= kRowsTabLds;
  switch (nbp) {
    EDLGT_RC_NB(1)
    EDLGT_RC_NB(2)
    EDLGT_RC_NB(4)
    EDLGT_RC_NB(8)
    EDLGT_RC_NB(16)
    default:
      return set_error(EDLGT_ERR_INVALID_ARG, "unsupported batch width");
  }
#undef EDLGT_RC_NB
#undef EDLGT_RC
  EDLGT_HIP(hipGetLastError());
  return EDLGT_OK;
}

template <class VT, class S, int G>
int launch_coef_nb(int nb, int nbp, int64_t n, const Csr& h, const VT* val, InCols<S> xs, OutCols<S> ys,
                   const double* coef, void* pack, hipStream_t st) {
  const unsigned pgrid = grid_for(n, 256, 65536);
  const int64_t blocks = std::min<int64_t>(65536, std::max<int64_t>(1, (n + 256 / G - 1) / (256 / G)));
  const uint8_t* term = h.n_terms > 1 ? h.term.as<uint8_t>() : (const uint8_t*)nullptr;
#define EDLGT_COEF_CASE(NBC)                                                                     \
  case NBC:                                                                                      \
    hipLaunchKernelGGL((pack_rows<S, NBC>), dim3(pgrid), dim3(256), 0, st, xs, nb, n, (S*)pack);   \
    hipLaunchKernelGGL((csr_spmm_coef_rows<VT, S, G, NBC>), dim3((unsigned)blocks), dim3(256), 0, st, \
                       n, h.row_ptr.as<int64_t>(), h.col.as<int32_t>(), val, term, (const S*)pack, \
                       ys, nb, coef);                                                            \
    break;
  switch (nbp) {
    EDLGT_COEF_CASE(1)
    EDLGT_COEF_CASE(2)
    EDLGT_COEF_CASE(4)
    EDLGT_COEF_CASE(8)
    EDLGT_COEF_CASE(16)
    default:
      return set_error(EDLGT_ERR_INVALID_ARG, "unsupported batch width");
  }
#undef EDLGT_COEF_CASE
  EDLGT_HIP(hipGetLastError());
  return EDLGT_OK;
}

template <class VT, class S>
int launch_coef(int g, int nb, int nbp, int64_t n, const Csr& h, const VT* val, InCols<S> xs, OutCols<S> ys,
                const double* coef, void* pack, hipStream_t st) {
  switch (g) {
    case 4: return launch_coef_nb<VT, S, 4>(nb, nbp, n, h, val, xs, ys, coef, pack, st);
    case 8: return launch_coef_nb<VT, S, 8>(nb, nbp, n, h, val, xs, ys, coef, pack, st);
    case 16: return launch_coef_nb<VT, S, 16>(nb, nbp, n, h, val, xs, ys, coef, pack, st);
    case 32: return launch_coef_nb<VT, S, 32>(nb, nbp, n, h, val, xs, ys, coef, pack, st);
  }
  return set_error(EDLGT_ERR_INVALID_ARG, "unsupported lanes-per-row");
}

}  // namespace

int coef_lanes(const Csr& h) {
  const double avg = h.n_rows > 0 ? (double)h.nnz / (double)h.n_rows : 1.0;
  return avg <= 6 ? 4 : avg <= 14 ? 8 : avg <= 40 ? 16 : 32;
}

int coef_pad(int nb) { return nb <= 1 ? 1 : nb <= 2 ? 2 : nb <= 4 ? 4 : nb <= 8 ? 8 : 16; }

// Whether the whole batch goes through one multi-RHS launch (row-interleaved
// right-hand sides): always for uncoded matrices; for coded ones while a full
// 16-wide packed block stays within a cache-sized footprint.  Larger coded
// matrices take groups of kGroupRhs columns per launch, zero-padded: a wider
// packed row would turn every gather into whole cache lines of a vector far
// larger than the MALL, while 4 columns still share one line per gather.
// Either way a column's arithmetic (kernel, lanes per row, summation order)
// depends on the matrix alone, never on the batch it is in.
constexpr size_t kFuseBytes = size_t(64) << 20;
constexpr int kGroupRhs = 4;
bool coef_fused(const Csr& h, int vec_type) {
  if (h.n_dict == 0) return true;
  const size_t es = vec_type == EDLGT_VALUE_C128 ? 16 : 8;
  return (size_t)h.n_rows * es * kCoefMaxRhs <= kFuseBytes;
}

static size_t coef_x_bytes(const Csr& h, int vec_type, int nb) {
  const int w = coef_fused(h, vec_type) ? coef_pad(nb) : kGroupRhs;
  const size_t b = (size_t)h.n_rows * w * (vec_type == EDLGT_VALUE_C128 ? 16 : 8);
  return (b + 255) / 256 * 256;
}

size_t coef_pack_bytes(const Csr& h, int vec_type, int nb) {
  return coef_x_bytes(h, vec_type, nb) + (h.n_dict > 0 ? (size_t)kCoefMaxRhs * h.n_dict * 16 : 0) + 256;
}

// xs / ys: device arrays of nb device pointers (hx / hy: the same pointers on
// the host); coef: device [nb][16] doubles (padded to the kernel width: unused
// rows may hold anything); pack: device workspace of coef_pack_bytes(h,
// vec_type, nb).  1 <= nb <= 16.  Vectors: complex128 for a complex matrix,
// float64 (F64) or complex128 (C128) for a real matrix.
int csr_spmm_coef(const Csr& h, int vec_type, int nb, const void* xs, void* ys, const void* const* hx,
                  void* const* hy, const double* coef, void* pack, hipStream_t st) {
  if (nb < 1 || nb > kCoefMaxRhs) return set_error(EDLGT_ERR_INVALID_ARG, "batch width must be in [1, 16]");
  const int g = coef_lanes(h);
  const int nbp = coef_pad(nb);
  const int64_t n = h.n_rows;
  const bool cplx = h.value_type == EDLGT_VALUE_C128;
  if (cplx && vec_type != EDLGT_VALUE_C128)
    return set_error(EDLGT_ERR_INVALID_ARG, "complex matrix needs complex vectors");
  if (h.n_dict > 0) {
    char* tab = static_cast<char*>(pack) + coef_x_bytes(h, vec_type, nb);
    const unsigned tg = grid_for((int64_t)nb * h.n_dict, 256, 1024);
    if (cplx)
      hipLaunchKernelGGL(scale_dict<double2>, dim3(tg), dim3(256), 0, st, h.dict_val.as<double2>(),
                         h.dict_term.as<int32_t>(), h.n_dict, coef, nb, (double2*)tab);
    else
      hipLaunchKernelGGL(scale_dict<double>, dim3(tg), dim3(256), 0, st, h.dict_val.as<double2>(),
                         h.dict_term.as<int32_t>(), h.n_dict, coef, nb, (double*)tab);
    EDLGT_HIP(hipGetLastError());
    const bool fused = coef_fused(h, vec_type);
    const int gw = fused ? nb : kGroupRhs;  // columns per launch
    const size_t te = cplx ? 16 : 8;
    for (int g0 = 0; g0 < nb; g0 += gw) {
      const int nbg = std::min(gw, nb - g0);
      const int wp = fused ? nbp : kGroupRhs;
      const char* tg = tab + (size_t)g0 * h.n_dict * te;
      const void* xg = static_cast<const char*>(xs) + (size_t)g0 * sizeof(void*);
      void* yg = static_cast<char*>(ys) + (size_t)g0 * sizeof(void*);
#define EDLGT_RCG(TT, S)                                                                          \
  {                                                                                               \
    const InCols<S> xc{(const S* const*)xg, nullptr, 0};                                          \
    const OutCols<S> yc{(S* const*)yg, nullptr, 0};                                               \
    const int ts = h.n_dict;                                                                      \
    switch (g) {                                                                                  \
      case 4: EDLGT_TRY((launch_rows_coded<TT, S, 4>(nbg, wp, h, (const TT*)tg, ts, xc, yc, pack, st))); break;   \
      case 8: EDLGT_TRY((launch_rows_coded<TT, S, 8>(nbg, wp, h, (const TT*)tg, ts, xc, yc, pack, st))); break;   \
      case 16: EDLGT_TRY((launch_rows_coded<TT, S, 16>(nbg, wp, h, (const TT*)tg, ts, xc, yc, pack, st))); break; \
      case 32: EDLGT_TRY((launch_rows_coded<TT, S, 32>(nbg, wp, h, (const TT*)tg, ts, xc, yc, pack, st))); break; \
      default: return set_error(EDLGT_ERR_INVALID_ARG, "unsupported lanes-per-row");              \
    }                                                                                             \
  }
      if (cplx) {
        EDLGT_RCG(double2, double2)
      } else if (vec_type == EDLGT_VALUE_F64) {
        EDLGT_RCG(double, double)
      } else {
        EDLGT_RCG(double, double2)
      }
#undef EDLGT_RCG
    }
    return EDLGT_OK;
  }
  if (cplx)
    return launch_coef<double2, double2>(g, nb, nbp, n, h, h.val.as<double2>(),
                                         InCols<double2>{(const double2* const*)xs, nullptr, 0},
                                         OutCols<double2>{(double2* const*)ys, nullptr, 0}, coef, pack, st);
  if (vec_type == EDLGT_VALUE_F64)
    return launch_coef<double, double>(g, nb, nbp, n, h, h.val.as<double>(),
                                       InCols<double>{(const double* const*)xs, nullptr, 0},
                                       OutCols<double>{(double* const*)ys, nullptr, 0}, coef, pack, st);
  return launch_coef<double, double2>(g, nb, nbp, n, h, h.val.as<double>(),
                                      InCols<double2>{(const double2* const*)xs, nullptr, 0},
                                      OutCols<double2>{(double2* const*)ys, nullptr, 0}, coef, pack, st);
}

// Block product Y = H X (FP64, 2 <= n_rhs) through row-interleaved groups of
// up to kGroupRhs columns: one gather per stored entry fetches that entry's x
// for the whole group.  Same kernels as the coefficient product with every
// coefficient 1, so a column's bits do not depend on the other columns.
size_t spmm_pack_bytes(const Csr& h, int xtype) {
  const size_t es = xtype == EDLGT_VALUE_C128 ? 16 : 8;
  return ((size_t)h.n_rows * kGroupRhs * es + 255) / 256 * 256;
}

bool spmm_packable(const Csr& h, int xtype) {
  if (h.mf) return false;
  if (h.value_type == EDLGT_VALUE_C128) return xtype == EDLGT_VALUE_C128;
  return xtype == EDLGT_VALUE_F64 || xtype == EDLGT_VALUE_C128;
}

int spmm_packed(const Csr& h, int xtype, int n_rhs, const void* x, int64_t ldx, void* y, int64_t ldy,
                void* pack, hipStream_t st) {
  const int g = coef_lanes(h);
  const int64_t n = h.n_rows;
  const bool cplx = h.value_type == EDLGT_VALUE_C128;
  const size_t es = xtype == EDLGT_VALUE_C128 ? 16 : 8;
  for (int g0 = 0; g0 < n_rhs; g0 += kGroupRhs) {
    const int nb = std::min(kGroupRhs, n_rhs - g0);
    const int nbp = coef_pad(nb);
    const char* xg = static_cast<const char*>(x) + (size_t)g0 * ldx * es;
    char* yg = static_cast<char*>(y) + (size_t)g0 * ldy * es;
#define EDLGT_PK(TT, S)                                                                           \
  {                                                                                               \
    const InCols<S> xc{nullptr, (const S*)xg, ldx};                                               \
    const OutCols<S> yc{nullptr, (S*)yg, ldy};                                                    \
    if (h.n_dict > 0) {                                                                           \
      const TT* tu = h.tab_unit.as<TT>();                                                         \
      switch (g) {                                                                                \
        case 4: EDLGT_TRY((launch_rows_coded<TT, S, 4>(nb, nbp, h, tu, 0, xc, yc, pack, st))); break;   \
        case 8: EDLGT_TRY((launch_rows_coded<TT, S, 8>(nb, nbp, h, tu, 0, xc, yc, pack, st))); break;   \
        case 16: EDLGT_TRY((launch_rows_coded<TT, S, 16>(nb, nbp, h, tu, 0, xc, yc, pack, st))); break; \
        case 32: EDLGT_TRY((launch_rows_coded<TT, S, 32>(nb, nbp, h, tu, 0, xc, yc, pack, st))); break; \
        default: return set_error(EDLGT_ERR_INVALID_ARG, "unsupported lanes-per-row");            \
      }                                                                                           \
    } else {                                                                                      \
      EDLGT_TRY((launch_coef<TT, S>(g, nb, nbp, n, h, h.val.as<TT>(), xc, yc, nullptr, pack, st))); \
    }                                                                                             \
  }
    if (cplx) {
      EDLGT_PK(double2, double2)
    } else if (xtype == EDLGT_VALUE_F64) {
      EDLGT_PK(double, double)
    } else {
      EDLGT_PK(double, double2)
    }
#undef EDLGT_PK
  }
  return EDLGT_OK;
}

int csr_autotune(Csr& h, int precision, int xtype, int n_rhs, const void* x, void* y,
                 hipStream_t st, int32_t* best_g, int64_t* best_grid, float* best_ms) {
  if (h.n_dict > 0 && h.stream_blocks.p) {
    // coded values: the launch variant of the coded stream kernel
    if (best_g) *best_g = 0;
    if (best_grid) *best_grid = -h.coded_variant - 1;
    EDLGT_TRY(coded_autotune(h, xtype, x, y, st, best_ms));
    if (best_grid) *best_grid = -h.coded_variant - 1;
    return EDLGT_OK;
  }
  const double avg = h.n_rows > 0 ? (double)h.nnz / (double)h.n_rows : 1.0;
  const int gs[] = {4, 8, 16, 32, 64};
  const int64_t grids[] = {4096, 16384, 65536, 262144};
  hipEvent_t e0, e1;
  EDLGT_HIP(hipEventCreate(&e0));
  EDLGT_HIP(hipEventCreate(&e1));
  const int32_t save_g = h.g_override;
  const int64_t save_grid = h.grid_override;
  float best = 1e30f;
  int32_t bg = 0;
  int64_t bgrid = 0;
  int status = EDLGT_OK;
  int best_plain = -1;
  if (n_rhs == 1 && h.stream_blocks.p) {  // CSR-stream candidates (lanes 0), every launch variant
    h.g_override = 0;
    h.grid_override = 0;
    const int save_v = h.plain_variant;
    for (int v = 0; v < kNumStreamVariants && status == EDLGT_OK; ++v) {
      status = csr_prepare_tiles(h, kStreamVariants[v].tile, st);
      if (status != EDLGT_OK) break;
      h.plain_variant = v;
      status = csr_spmm(h, precision, xtype, n_rhs, x, h.n_rows, y, h.n_rows, st);  // warm
      if (status != EDLGT_OK) break;
      (void)hipEventRecord(e0, st);
      for (int r = 0; r < 3 && status == EDLGT_OK; ++r)
        status = csr_spmm(h, precision, xtype, n_rhs, x, h.n_rows, y, h.n_rows, st);
      (void)hipEventRecord(e1, st);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms / 3 < best) {
        best = ms / 3;
        best_plain = v;
      }
    }
    h.plain_variant = best_plain >= 0 ? best_plain : save_v;
  }
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
  // the stream kernel won: its launch variant is reported as grid = -(variant + 1)
  if (best_grid) *best_grid = (bg == 0 && best_plain >= 0) ? -(int64_t)h.plain_variant - 1 : bgrid;
  if (best_ms) *best_ms = best;
  return EDLGT_OK;
}

}  // namespace edlgt
