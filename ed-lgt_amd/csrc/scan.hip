// Device-wide exclusive prefix sums (int32 or int64 counts -> int64 offsets).
//
// Three phases per level: (1) every 256-thread block reduces a tile of
// 256*ITEMS counts, (2) the block sums are scanned (recursively), (3) every
// block rescans its tile in LDS and adds its block offset.  offsets[n] holds
// the total.  Used for the per-prefix completion counts of sector enumeration
// and the per-row nonzero counts of CSR assembly.
#include "common.h"

namespace edlgt {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;

__device__ inline int64_t block_exclusive_scan(int64_t v, int64_t* lds, int64_t* total) {
  // Wave-level inclusive scan with shuffles, then across the 4 waves.
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  int64_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wave] = x;
  __syncthreads();
  int64_t wave_off = 0, tot = 0;
  for (int w = 0; w < kBlock / 64; ++w) {
    int64_t s = lds[w];
    if (w < wave) wave_off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return wave_off + x - v;
}

template <class T>
__global__ void __launch_bounds__(kBlock) tile_reduce(const T* __restrict__ in, int64_t n,
                                                      int64_t* __restrict__ block_sums) {
  __shared__ int64_t lds[kBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kTile;
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    int64_t idx = base + (int64_t)i * kBlock + threadIdx.x;
    if (idx < n) s += (int64_t)in[idx];
  }
  int64_t tot;
  (void)block_exclusive_scan(s, lds, &tot);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

template <class T>
__global__ void __launch_bounds__(kBlock) tile_scan(const T* __restrict__ in, int64_t n,
                                                    const int64_t* __restrict__ block_offsets,
                                                    int64_t* __restrict__ out) {
  __shared__ int64_t lds[kBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
  int64_t v[kItems];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    int64_t idx = base + i;
    v[i] = idx < n ? (int64_t)in[idx] : 0;
    s += v[i];
  }
  int64_t tot;
  int64_t run = block_exclusive_scan(s, lds, &tot) + block_offsets[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    int64_t idx = base + i;
    if (idx < n) out[idx] = run;
    run += v[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) out[n] = block_offsets[blockIdx.x] + tot;
}

template <class T>
int scan_impl(const T* counts, int64_t* offsets, int64_t n, hipStream_t s) {
  if (n <= 0) {
    EDLGT_HIP(hipMemsetAsync(offsets, 0, sizeof(int64_t), s));
    return EDLGT_OK;
  }
  const int64_t nb = (n + kTile - 1) / kTile;
  if (nb > 0x7fffffffLL) return set_error(EDLGT_ERR_INDEX_OVERFLOW, "scan too large");
  DevBuf sums, sum_off;
  EDLGT_TRY(sums.alloc(nb * sizeof(int64_t)));
  EDLGT_TRY(sum_off.alloc((nb + 1) * sizeof(int64_t)));
  hipLaunchKernelGGL(tile_reduce<T>, dim3((unsigned)nb), dim3(kBlock), 0, s, counts, n,
                     sums.as<int64_t>());
  EDLGT_HIP(hipGetLastError());
  if (nb == 1) {
    EDLGT_HIP(hipMemsetAsync(sum_off.p, 0, sizeof(int64_t), s));
  } else {
    EDLGT_TRY(scan_impl<int64_t>(sums.as<int64_t>(), sum_off.as<int64_t>(), nb, s));
  }
  hipLaunchKernelGGL(tile_scan<T>, dim3((unsigned)nb), dim3(kBlock), 0, s, counts, n,
                     sum_off.as<int64_t>(), offsets);
  EDLGT_HIP(hipGetLastError());
  // The temporaries are released on return: wait for the kernels using them.
  EDLGT_HIP(hipStreamSynchronize(s));
  return EDLGT_OK;
}

}  // namespace

int exclusive_scan_i32_i64(const int32_t* counts, int64_t* offsets, int64_t n, hipStream_t s) {
  return scan_impl<int32_t>(counts, offsets, n, s);
}
int exclusive_scan_i64(const int64_t* counts, int64_t* offsets, int64_t n, hipStream_t s) {
  return scan_impl<int64_t>(counts, offsets, n, s);
}

}  // namespace edlgt
