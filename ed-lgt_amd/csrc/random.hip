// Uniform [0, 1) doubles from a PCG64 stream on the device: the generator
// behind numpy.random.default_rng (128-bit LCG s <- a s + inc, XSL-RR 64-bit
// output, value = (out >> 11) * 2^-53).  The start vector of the model
// layer's eigensolves is default_rng(0).random(n); generating it here gives
// the same numbers bit for bit without a host pass and a PCIe copy of a
// full-length vector (C5: 462 MB as complex128).
#include "common.h"
#include "kernels.h"

namespace edlgt {
namespace {

using u128 = unsigned __int128;

// PCG64's default 128-bit multiplier (O'Neill's PCG family)
constexpr u128 kPcgMult = ((u128)0x2360ED051FC65DA4ull << 64) | (u128)0x4385DF649FCCF645ull;
constexpr int kTB = 256;
constexpr int kPerThread = 16;

__host__ __device__ inline uint64_t pcg_output(u128 s) {
  const uint64_t v = (uint64_t)(s >> 64) ^ (uint64_t)s;
  const unsigned r = (unsigned)(s >> 122);
  return (v >> r) | (v << ((64u - r) & 63u));
}

// The affine map of `delta` LCG steps, s -> m s + p (Brown, "Random number
// generation with arbitrary strides", 1994: square-and-multiply on the map).
struct Affine {
  u128 m, p;
};
__host__ __device__ inline Affine lcg_jump(u128 mult, u128 inc, uint64_t delta) {
  u128 am = 1, ap = 0;
  while (delta) {
    if (delta & 1) {
      am *= mult;
      ap = ap * mult + inc;
    }
    inc = (mult + 1) * inc;
    mult *= mult;
    delta >>= 1;
  }
  return Affine{am, ap};
}

__device__ inline void put(double* out, int64_t i, double v) { out[i] = v; }
__device__ inline void put(double2* out, int64_t i, double v) { out[i] = make_double2(v, 0.0); }

// Value i comes from state s_{i+1}.  Thread t of block b writes entries
// b*kTB*kPerThread + j*kTB + t (coalesced stores), stepping its state by the
// precomputed kTB-step map between them.
template <class S>
__global__ void __launch_bounds__(kTB) pcg64_uniform(S* __restrict__ out, int64_t n, u128 s0, u128 inc,
                                                     u128 step_m, u128 step_p) {
  const int64_t i0 = (int64_t)blockIdx.x * kTB * kPerThread + threadIdx.x;
  if (i0 >= n) return;
  const Affine a = lcg_jump(kPcgMult, inc, (uint64_t)i0 + 1);
  u128 s = a.m * s0 + a.p;
#pragma unroll 4
  for (int j = 0; j < kPerThread; ++j) {
    const int64_t i = i0 + (int64_t)j * kTB;
    if (i >= n) break;
    put(out, i, (double)(pcg_output(s) >> 11) * (1.0 / 9007199254740992.0));
    s = step_m * s + step_p;
  }
}

}  // namespace

int random_uniform(void* out, int64_t n, bool complex_values, const uint64_t state[4], hipStream_t st) {
  if (n <= 0) return EDLGT_OK;
  const u128 s0 = ((u128)state[0] << 64) | (u128)state[1];
  const u128 inc = ((u128)state[2] << 64) | (u128)state[3];
  const Affine step = lcg_jump(kPcgMult, inc, (uint64_t)kTB);
  const int64_t per_block = (int64_t)kTB * kPerThread;
  const unsigned grid = (unsigned)((n + per_block - 1) / per_block);
  if (complex_values)
    hipLaunchKernelGGL(pcg64_uniform<double2>, dim3(grid), dim3(kTB), 0, st, (double2*)out, n, s0, inc, step.m,
                       step.p);
  else
    hipLaunchKernelGGL(pcg64_uniform<double>, dim3(grid), dim3(kTB), 0, st, (double*)out, n, s0, inc, step.m,
                       step.p);
  EDLGT_HIP(hipGetLastError());
  return EDLGT_OK;
}

}  // namespace edlgt
