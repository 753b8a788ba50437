// Tall-skinny Gram products G = V^H W of the Krylov solvers (complex128 or
// float64 columns, column-major with leading dimension ld) and the device
// scalar helpers they share with the other solver kernels.  Shared by
// lanczos.hip and the Gram benchmark harness (ed-lgt_amd/perf).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace edlgt {
namespace gramk {

// ---------------------------------------------------------------- scalars
__device__ inline double2 cj_mul(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ inline double cj_mul(double a, double b) { return a * b; }
__device__ inline double2 add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double add(double a, double b) { return a + b; }
__device__ inline double2 mul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// V element times a complex coefficient, in the vector's scalar type (real
// vectors use the real part: their coefficients are real).
__device__ inline double2 mul_s(double2 v, double2 c) { return mul(v, c); }
__device__ inline double mul_s(double v, double2 c) { return v * c.x; }
template <class S> __device__ inline S zero();
template <> __device__ inline double zero<double>() { return 0.0; }
template <> __device__ inline double2 zero<double2>() { return make_double2(0.0, 0.0); }
__device__ inline double2 to_c(double2 a) { return a; }
__device__ inline double2 to_c(double a) { return make_double2(a, 0.0); }

constexpr int kGramBlocks = 1024;
constexpr int kGramCols = 64;  // V columns per Gram launch (one per lane)
constexpr int kGramRows = 8;   // rows per thread per sweep
constexpr int kTB = 256;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double2 wave_sum(double2 v) { return make_double2(wave_sum(v.x), wave_sum(v.y)); }

// partial[blk][j][c] = sum over the block's rows of conj(V_j[r]) W_c[r],
// j < nv <= 64, c < nw <= NW.  Each thread keeps kGramRows rows of W in
// registers for a sweep, so W and V are each read once; for every column j
// the lane products are summed across the wave with a fixed xor butterfly
// and lane j keeps the running sum of column j.  The four waves' sums are
// combined in fixed order: the result is bitwise reproducible.
template <class S, int NW>
__global__ void __launch_bounds__(kTB) gram_partial(const S* __restrict__ V, int64_t ldv, int nv,
                                                    const S* __restrict__ W, int64_t ldw, int nw,
                                                    int64_t n, double2* __restrict__ partial) {
  __shared__ double2 red[kTB / 64][kGramCols][NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)blockIdx.x * chunk;
  const int64_t r1 = min(n, r0 + chunk);
  S mine[NW];
#pragma unroll
  for (int c = 0; c < NW; ++c) mine[c] = zero<S>();
  for (int64_t base = r0; base < r1; base += (int64_t)kTB * kGramRows) {
    S w[kGramRows][NW];
#pragma unroll
    for (int i = 0; i < kGramRows; ++i) {
      const int64_t r = base + (int64_t)i * kTB + threadIdx.x;
#pragma unroll
      for (int c = 0; c < NW; ++c) w[i][c] = (r < r1 && c < nw) ? W[c * ldw + r] : zero<S>();
    }
    for (int j = 0; j < nv; ++j) {
      S acc[NW];
#pragma unroll
      for (int c = 0; c < NW; ++c) acc[c] = zero<S>();
#pragma unroll
      for (int i = 0; i < kGramRows; ++i) {
        const int64_t r = base + (int64_t)i * kTB + threadIdx.x;
        const S v = r < r1 ? V[(int64_t)j * ldv + r] : zero<S>();
#pragma unroll
        for (int c = 0; c < NW; ++c) acc[c] = add(acc[c], cj_mul(v, w[i][c]));
      }
#pragma unroll
      for (int c = 0; c < NW; ++c) {
        const S t = wave_sum(acc[c]);
        if (lane == j) mine[c] = add(mine[c], t);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < NW; ++c) red[wave][lane][c] = to_c(mine[c]);
  __syncthreads();
  for (int t = threadIdx.x; t < nv * NW; t += kTB) {
    const int j = t / NW, c = t % NW;
    double2 s = red[0][j][c];
#pragma unroll
    for (int q = 1; q < kTB / 64; ++q) s = add(s, red[q][j][c]);
    partial[((int64_t)blockIdx.x * nv + j) * NW + c] = s;
  }
}

// out[(j0 + j) * nw_real + c] = sum over blocks of partial[b][j][c]: one wave
// per output element, lanes over blocks, fixed butterfly.
template <int NW>
__global__ void __launch_bounds__(kTB) gram_finish(const double2* __restrict__ partial, int nblk, int nv,
                                                   int nw_real, int j0, double2* __restrict__ out) {
  const int e = blockIdx.x * (kTB / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (e >= nv * NW) return;  // whole waves exit together
  const int j = e / NW, c = e % NW;
  double2 s = make_double2(0.0, 0.0);
  for (int b = lane; b < nblk; b += 64) s = add(s, partial[((int64_t)b * nv + j) * NW + c]);
  s = wave_sum(s);
  if (lane == 0 && c < nw_real) out[(int64_t)(j0 + j) * nw_real + c] = s;
}

}  // namespace gramk
}  // namespace edlgt
