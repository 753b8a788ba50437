// Benchmark harness for the tall-skinny Gram product G = V^H W (complex128)
// of the Krylov solvers: the solvers' VALU kernel (csrc/gram.h) against two
// MFMA formulations (v_mfma_f64_16x16x4f64) and a register-accumulating VALU
// variant, on synthetic columns of a given length.  Not part of the library:
// built and run by gram_bench.py.
//
// MFMA formulation: rows are the reduction (K) dimension, one MFMA step takes
// 4 rows; A = 16 basis columns x 4 rows (lane l: column l % 16, row l / 16),
// B = [Re W | Im W] as 2 nw real columns (lane l: row l / 16, column l % 16),
// so one complex product is 2 real MFMAs: P1 = Re(V)^T B, P2 = Im(V)^T B,
// Re G = P1[:, c] + P2[:, nw + c], Im G = P1[:, nw + c] - P2[:, c].  The
// accumulator layout (checked by mfma_layout) is lane l -> D[4 q + l / 16][l % 16].
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../csrc/gram.h"

using namespace edlgt::gramk;
typedef double double4_t __attribute__((ext_vector_type(4)));

// ---- lane layout check of v_mfma_f64_16x16x4f64: A[i][k] = 1000 i + k,
// B[k][j] = (k + 1) * (j + 1) * 0.5; lane l supplies A[l%16][l/16], B[l/16][l%16]
__global__ void mfma_layout(double* out) {
  const int l = threadIdx.x;
  const double a = 1000.0 * (l % 16) + (l / 16);
  const double b = 0.5 * ((l / 16) + 1) * ((l % 16) + 1);
  double4_t c = {0, 0, 0, 0};
  c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  for (int v = 0; v < 4; ++v) out[l * 4 + v] = c[v];
}

// ---- MFMA variant.  Rows are the reduction (K) dimension: one MFMA step
// consumes 4 rows.  A = V^T tile (16 basis columns x 4 rows; lane l holds
// column l%16 at row l/16), B = [Re W | Im W] (4 rows x 2NW columns, zero
// beyond; lane l holds row l/16, column l%16).  P1 = Re(V)^T B, P2 = Im(V)^T B:
// G_re = P1[:, c] + P2[:, NW + c], G_im = P1[:, NW + c] - P2[:, c].
// The four waves of a block take consecutive 4-row steps of the block's row
// chunk.  partial[blk][j][c] as the VALU kernel (nv <= 16 * NG).
template <int NG, int NW>
__global__ void __launch_bounds__(kTB) gram_mfma(const double2* __restrict__ V, int64_t ldv, int nv,
                                                 const double2* __restrict__ W, int64_t ldw, int nw,
                                                 int64_t n, double2* __restrict__ partial) {
  static_assert(2 * NW <= 16, "W block too wide");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)blockIdx.x * chunk;
  const int64_t r1 = min(n, r0 + chunk);
  const int i16 = lane & 15, k4 = lane >> 4;
  double4_t p1[NG], p2[NG];
#pragma unroll
  for (int gi = 0; gi < NG; ++gi) { p1[gi] = (double4_t){0, 0, 0, 0}; p2[gi] = (double4_t){0, 0, 0, 0}; }
  // B operand column: i16 < NW -> Re W[:, i16]; NW <= i16 < 2NW -> Im W[:, i16 - NW]
  const int wc = i16 < NW ? i16 : i16 - NW;
  const bool wre = i16 < NW, wok = i16 < 2 * NW && wc < nw;
  for (int64_t base = r0 + 4 * wave; base < r1; base += 16) {
    const int64_t r = base + k4;
    const bool rok = r < r1;
    double b = 0.0;
    if (wok && rok) {
      const double2 w = W[(int64_t)wc * ldw + r];
      b = wre ? w.x : w.y;
    }
#pragma unroll
    for (int gi = 0; gi < NG; ++gi) {
      const int j = gi * 16 + i16;
      double2 v = make_double2(0, 0);
      if (rok && j < nv) v = V[(int64_t)j * ldv + r];
      p1[gi] = __builtin_amdgcn_mfma_f64_16x16x4f64(v.x, b, p1[gi], 0, 0, 0);
      p2[gi] = __builtin_amdgcn_mfma_f64_16x16x4f64(v.y, b, p2[gi], 0, 0, 0);
    }
  }
  // lane l holds P[4 q + l/16][l % 16], q < 4
  __shared__ double red[kTB / 64][NG][2][16][16];
#pragma unroll
  for (int gi = 0; gi < NG; ++gi)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      red[wave][gi][0][4 * q + k4][i16] = p1[gi][q];
      red[wave][gi][1][4 * q + k4][i16] = p2[gi][q];
    }
  __syncthreads();
  for (int t = threadIdx.x; t < nv * NW; t += kTB) {
    const int j = t / NW, c = t % NW, gi = j / 16, jj = j % 16;
    double gr = 0.0, gim = 0.0;
#pragma unroll
    for (int q = 0; q < kTB / 64; ++q) {
      gr += red[q][gi][0][jj][c] + red[q][gi][1][jj][NW + c];
      gim += red[q][gi][0][jj][NW + c] - red[q][gi][1][jj][c];
    }
    partial[((int64_t)blockIdx.x * nv + j) * NW + c] = make_double2(gr, gim);
  }
}


// ---- MFMA with LDS staging: per wave, tiles of 32 rows; a 16-column group of
// V is loaded coalesced (lane = row, 2 columns per instruction) into the
// wave's LDS slice [col][row] (row stride 16 B, column stride 528 B), then
// the 8 MFMA k-steps of the tile read their A operand from LDS.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <int NG, int NW>
__global__ void __launch_bounds__(kTB) gram_mfma_lds(const double2* __restrict__ V, int64_t ldv, int nv,
                                                     const double2* __restrict__ W, int64_t ldw, int nw,
                                                     int64_t n, double2* __restrict__ partial) {
  static_assert(2 * NW <= 16, "W block too wide");
  constexpr int TR = 32, CS = TR + 1;  // rows per tile, padded column stride (elements)
  __shared__ double2 tile_s[kTB / 64][16 * CS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double2* tl = tile_s[wave];
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)blockIdx.x * chunk;
  const int64_t r1 = min(n, r0 + chunk);
  const int i16 = lane & 15, k4 = lane >> 4;
  double4_t p1[NG], p2[NG];
#pragma unroll
  for (int gi = 0; gi < NG; ++gi) { p1[gi] = (double4_t){0, 0, 0, 0}; p2[gi] = (double4_t){0, 0, 0, 0}; }
  const int wc = i16 < NW ? i16 : i16 - NW;
  const bool wre = i16 < NW, wok = i16 < 2 * NW && wc < nw;
  const int lrow = lane & 31, lcol = lane >> 5;
  for (int64_t base = r0 + (int64_t)TR * wave; base < r1; base += (int64_t)TR * (kTB / 64)) {
    double b[TR / 4];
#pragma unroll
    for (int s = 0; s < TR / 4; ++s) {
      const int64_t r = base + 4 * s + k4;
      b[s] = 0.0;
      if (wok && r < r1) {
        const double2 w = W[(int64_t)wc * ldw + r];
        b[s] = wre ? w.x : w.y;
      }
    }
#pragma unroll
    for (int gi = 0; gi < NG; ++gi) {
      const int64_t r = base + lrow;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int col = 2 * c + lcol, j = gi * 16 + col;
        tl[col * CS + lrow] = (r < r1 && j < nv) ? V[(int64_t)j * ldv + r] : make_double2(0, 0);
      }
      wave_sync();
#pragma unroll
      for (int s = 0; s < TR / 4; ++s) {
        const double2 v = tl[i16 * CS + 4 * s + k4];
        p1[gi] = __builtin_amdgcn_mfma_f64_16x16x4f64(v.x, b[s], p1[gi], 0, 0, 0);
        p2[gi] = __builtin_amdgcn_mfma_f64_16x16x4f64(v.y, b[s], p2[gi], 0, 0, 0);
      }
      wave_sync();
    }
  }
  __syncthreads();
  double* red = reinterpret_cast<double*>(tile_s);  // [wave][2][16][16] per group, reused
  static_assert(sizeof(tile_s) >= sizeof(double) * (kTB / 64) * 2 * 256, "reduction buffer");
  for (int gi = 0; gi < NG; ++gi) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      red[((wave * 2 + 0) * 16 + 4 * q + k4) * 16 + i16] = p1[gi][q];
      red[((wave * 2 + 1) * 16 + 4 * q + k4) * 16 + i16] = p2[gi][q];
    }
    __syncthreads();
    const int ncol = min(16, nv - gi * 16);
    for (int t = threadIdx.x; t < ncol * NW; t += kTB) {
      const int jj = t / NW, c = t % NW;
      double gr = 0.0, gim = 0.0;
#pragma unroll
      for (int q = 0; q < kTB / 64; ++q) {
        gr += red[((q * 2 + 0) * 16 + jj) * 16 + c] + red[((q * 2 + 1) * 16 + jj) * 16 + NW + c];
        gim += red[((q * 2 + 0) * 16 + jj) * 16 + NW + c] - red[((q * 2 + 1) * 16 + jj) * 16 + c];
      }
      partial[((int64_t)blockIdx.x * nv + gi * 16 + jj) * NW + c] = make_double2(gr, gim);
    }
    __syncthreads();
  }
}

// ---- VALU variant without per-column butterflies: each thread keeps its own
// sums for every (column, w) pair over its rows; one block reduction at the end.
template <int NV, int NW>
__global__ void __launch_bounds__(kTB) gram_acc(const double2* __restrict__ V, int64_t ldv, int nv,
                                                const double2* __restrict__ W, int64_t ldw, int nw,
                                                int64_t n, double2* __restrict__ partial) {
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)blockIdx.x * chunk;
  const int64_t r1 = min(n, r0 + chunk);
  double2 acc[NV][NW];
#pragma unroll
  for (int j = 0; j < NV; ++j)
#pragma unroll
    for (int c = 0; c < NW; ++c) acc[j][c] = make_double2(0, 0);
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kTB) {
    double2 w[NW];
#pragma unroll
    for (int c = 0; c < NW; ++c) w[c] = c < nw ? W[c * ldw + r] : make_double2(0, 0);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const double2 v = j < nv ? V[(int64_t)j * ldv + r] : make_double2(0, 0);
#pragma unroll
      for (int c = 0; c < NW; ++c) acc[j][c] = add(acc[j][c], cj_mul(v, w[c]));
    }
  }
  __shared__ double2 red[kTB / 64][NV][NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; ++j)
#pragma unroll
    for (int c = 0; c < NW; ++c) {
      const double2 t = wave_sum(acc[j][c]);
      if (lane == 0) red[wave][j][c] = t;
    }
  __syncthreads();
  for (int t = threadIdx.x; t < nv * NW; t += kTB) {
    const int j = t / NW, c = t % NW;
    double2 s = red[0][j][c];
#pragma unroll
    for (int q = 1; q < kTB / 64; ++q) s = add(s, red[q][j][c]);
    partial[((int64_t)blockIdx.x * nv + j) * NW + c] = s;
  }
}

__global__ void fill(double2* x, int64_t n, uint64_t seed) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t h = (seed + i) * 0x9e3779b97f4a7c15ull; h ^= h >> 31; h *= 0xbf58476d1ce4e5b9ull; h ^= h >> 29;
    x[i] = make_double2((double)(h >> 11) * 0x1p-53 - 0.5, (double)((h * 7) >> 11) * 0x1p-53 - 0.5);
  }
}

extern "C" {
int probe_layout(double* host_out) {
  double* d;
  if (hipMalloc(&d, 64 * 4 * sizeof(double)) != hipSuccess) return -1;
  hipLaunchKernelGGL(mfma_layout, dim3(1), dim3(64), 0, 0, d);
  if (hipMemcpy(host_out, d, 64 * 4 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -2;
  (void)hipFree(d);
  return 0;
}

// kind 0: current VALU kernel; 1: MFMA; 2: register-accumulating VALU.
// V: n x nv, W: n x nw (device, complex128, column-major); out: nv x nw host.
int probe_gram(int kind, int64_t n, int nv, int nw, int nblk, int reps, double* out_host, float* ms) {
  if (nv < 1 || nv > 64 || nw < 1 || nw > 8 || (kind == 2 && nv > 16) || nblk > 4096) return -9;
  double2 *V, *W, *part, *out;
  if (hipMalloc(&V, (size_t)n * nv * 16) || hipMalloc(&W, (size_t)n * nw * 16) ||
      hipMalloc(&part, (size_t)nblk * 64 * 8 * 16) || hipMalloc(&out, 64 * 8 * 16)) return -1;
  hipLaunchKernelGGL(fill, dim3(4096), dim3(256), 0, 0, V, n * nv, 1);
  hipLaunchKernelGGL(fill, dim3(4096), dim3(256), 0, 0, W, n * nw, 777);
  const int nwp = nw <= 1 ? 1 : nw <= 2 ? 2 : nw <= 4 ? 4 : 8;
  auto launch = [&]() {
#define CASEW(NWC)                                                                                      \
    if (nwp == NWC) {                                                                                   \
      if (kind == 0) hipLaunchKernelGGL((gram_partial<double2, NWC>), dim3(nblk), dim3(kTB), 0, 0, V, n, nv, W, n, nw, n, part); \
      if (kind == 1) {                                                                                  \
        if (nv <= 16) hipLaunchKernelGGL((gram_mfma<1, NWC>), dim3(nblk), dim3(kTB), 0, 0, V, n, nv, W, n, nw, n, part); \
        else if (nv <= 32) hipLaunchKernelGGL((gram_mfma<2, NWC>), dim3(nblk), dim3(kTB), 0, 0, V, n, nv, W, n, nw, n, part); \
        else hipLaunchKernelGGL((gram_mfma<4, NWC>), dim3(nblk), dim3(kTB), 0, 0, V, n, nv, W, n, nw, n, part); \
      }                                                                                                 \
      if (kind == 3) {                                                                                  \
        if (nv <= 16) hipLaunchKernelGGL((gram_mfma_lds<1, NWC>), dim3(nblk), dim3(kTB), 0, 0, V, n, nv, W, n, nw, n, part); \
        else if (nv <= 32) hipLaunchKernelGGL((gram_mfma_lds<2, NWC>), dim3(nblk), dim3(kTB), 0, 0, V, n, nv, W, n, nw, n, part); \
        else hipLaunchKernelGGL((gram_mfma_lds<4, NWC>), dim3(nblk), dim3(kTB), 0, 0, V, n, nv, W, n, nw, n, part); \
      }                                                                                                 \
      if (kind == 2) {                                                                                  \
        if (nv <= 8) hipLaunchKernelGGL((gram_acc<8, NWC>), dim3(nblk), dim3(kTB), 0, 0, V, n, nv, W, n, nw, n, part); \
        else hipLaunchKernelGGL((gram_acc<16, NWC>), dim3(nblk), dim3(kTB), 0, 0, V, n, nv, W, n, nw, n, part); \
      }                                                                                                 \
      hipLaunchKernelGGL((gram_finish<NWC>), dim3((nv * NWC + 3) / 4), dim3(kTB), 0, 0, part, nblk, nv, nw, 0, out); \
    }
    CASEW(1) CASEW(2) CASEW(4) CASEW(8)
#undef CASEW
  };
  launch();
  if (hipDeviceSynchronize() != hipSuccess) return -2;
  hipEvent_t a, b;
  (void)hipEventCreate(&a); (void)hipEventCreate(&b);
  (void)hipEventRecord(a, 0);
  for (int i = 0; i < reps; ++i) launch();
  (void)hipEventRecord(b, 0);
  (void)hipEventSynchronize(b);
  (void)hipEventElapsedTime(ms, a, b);
  *ms /= reps;
  (void)hipMemcpy(out_host, out, (size_t)nv * nw * 16, hipMemcpyDeviceToHost);
  (void)hipFree(V); (void)hipFree(W); (void)hipFree(part); (void)hipFree(out);
  (void)hipEventDestroy(a); (void)hipEventDestroy(b);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}
}
