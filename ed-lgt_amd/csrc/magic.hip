// Stabilizer Renyi-2 entropy of a sector state on a truncated support
// (QMB_state.stabilizer_renyi_entropy, edlgt/modeling/qmb_state.py:393-465,
// with extract_support :708-790 and the kernels of edlgt/tools/magic.py).
//
//   support  the fewest largest-|c|^2 states holding >= 1 - delta of the
//            weight (probabilities sorted descending, stable; the cumulative
//            sum taken left to right on the host as numpy's cumsum does),
//            ordered by key, coefficients renormalised over the support;
//   strings  X-strings p = (beta - alpha) mod d of every ordered pair of
//            support states, keys with the rightmost site fastest: one
//            radix sort of the K^2 pair keys groups the pairs by string, and
//            a group is exactly the list {(alpha, alpha + p)} of the string's
//            nonzero A_alpha = c_alpha conj(c_{alpha+p});
//   T_p      sum_kappa |V_kappa|^2, V_kappa = sum over ordered pairs (a, b) of
//            the string's list with alpha_a + alpha_b = kappa (mod d) of
//            A_a conj(A_b): the n_A^2 pairs of every string are generated in
//            batches of whole strings (one thread per pair), sorted by kappa
//            and then (stably) by string, and each (string, kappa) run is
//            summed in its sorted order;
//   M2       sum of |V|^2 over all runs, in run order (fixed-order block
//            partials, summed on the host in block order): SRE2 = -log M2.
//
// Every sort is a stable radix sort, so the result is deterministic.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>

#include "common.h"
#include "kernels.h"

namespace edlgt {
namespace {

constexpr int kTB = 256;
constexpr int64_t kPairBudget = int64_t(1) << 27;  // kappa pairs per batch
constexpr int kXBits = 14;                          // support index bits in a pair code

inline unsigned grid_of(int64_t n, int64_t cap = 65535) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kTB - 1) / kTB, cap));
}

__global__ void prob_iota(const double2* __restrict__ psi, int64_t n, double* __restrict__ prob,
                          uint32_t* __restrict__ idx) {
  for (int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTB) {
    const double2 c = psi[i];
    prob[i] = c.x * c.x + c.y * c.y;
    idx[i] = (uint32_t)i;
  }
}

// Support rows: coefficient, configuration digits (int32, n_sites per row)
// and key of the states listed in sidx (ascending = key order).
__global__ void gather_support(const uint32_t* __restrict__ sidx, int64_t K, const double2* __restrict__ psi,
                               const uint8_t* __restrict__ cfg, int cfg_stride, int n_sites,
                               double2* __restrict__ c, int32_t* __restrict__ scfg) {
  for (int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x; t < K; t += (int64_t)gridDim.x * kTB) {
    const int64_t r = sidx[t];
    c[t] = psi[r];
    for (int s = 0; s < n_sites; ++s) scfg[t * n_sites + s] = cfg[r * cfg_stride + s];
  }
}

__device__ inline uint64_t key_of_sum(const int32_t* __restrict__ a, const int32_t* __restrict__ b, int sign,
                                      const int32_t* __restrict__ dims, const uint64_t* __restrict__ strides,
                                      int n_sites) {
  uint64_t k = 0;
  for (int s = 0; s < n_sites; ++s) {
    const int32_t d = dims[s];
    int32_t v = (a[s] + sign * b[s]) % d;
    if (v < 0) v += d;
    k += (uint64_t)v * strides[s];
  }
  return k;
}

// Pair q = i * K + j: the X-string taking support state i to state j.
__global__ void string_keys(const int32_t* __restrict__ scfg, int64_t K, int n_sites,
                            const int32_t* __restrict__ dims, const uint64_t* __restrict__ strides,
                            uint64_t* __restrict__ key, uint32_t* __restrict__ pay) {
  const int64_t n = K * K;
  for (int64_t q = (int64_t)blockIdx.x * kTB + threadIdx.x; q < n; q += (int64_t)gridDim.x * kTB) {
    const int64_t i = q / K, j = q - i * K;
    key[q] = key_of_sum(scfg + j * n_sites, scfg + i * n_sites, -1, dims, strides, n_sites);
    pay[q] = (uint32_t)q;
  }
}

__global__ void head_flags(const uint64_t* __restrict__ key, int64_t n, int32_t* __restrict__ flag) {
  for (int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x; t < n; t += (int64_t)gridDim.x * kTB)
    flag[t] = (t == 0 || key[t] != key[t - 1]) ? 1 : 0;
}

// starts[pos[t]] = t for every head t; starts[S] = n.
__global__ void scatter_heads(const int32_t* __restrict__ flag, const int64_t* __restrict__ pos, int64_t n,
                              int64_t* __restrict__ starts) {
  for (int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x; t < n; t += (int64_t)gridDim.x * kTB)
    if (flag[t]) starts[pos[t]] = t;
  if (blockIdx.x == 0 && threadIdx.x == 0) starts[pos[n]] = n;
}

__global__ void seg_work(const int64_t* __restrict__ seg_start, int64_t S, int64_t* __restrict__ work) {
  for (int64_t s = (int64_t)blockIdx.x * kTB + threadIdx.x; s < S; s += (int64_t)gridDim.x * kTB) {
    const int64_t n = seg_start[s + 1] - seg_start[s];
    work[s] = n * n;
  }
}

struct PairCtx {
  const int64_t* seg_start;  // [S + 1] string groups in the sorted pair list
  const uint32_t* spay;      // sorted pair payloads q = i * K + j
  const double2* c;          // support coefficients (normalised)
  const int32_t* scfg;       // support digits
  const int32_t* dims;
  const uint64_t* strides;
  int64_t K;
  int n_sites;
};

__device__ inline void pair_member(const PairCtx& P, int64_t s, int64_t x, int64_t* i_alpha, double2* A) {
  const uint32_t q = P.spay[P.seg_start[s] + x];
  const int64_t i = q / P.K, j = q - i * P.K;
  const double2 ci = P.c[i], cj = P.c[j];
  // A = c_i conj(c_j)
  *A = make_double2(ci.x * cj.x + ci.y * cj.y, ci.y * cj.x - ci.x * cj.y);
  *i_alpha = i;
}

// Pair q of the batch [s0, s1) (q < wp[s1] - wp[s0]): its string s (binary
// search in the batch's work prefix wp), and within it the ordered pair
// (x, y) = divmod(q - (wp[s] - wp[s0]), n_A); key kappa and code
// (s - s0) << 2 kXBits | x << kXBits | y.
__global__ void kappa_pairs(PairCtx P, int64_t s0, int64_t s1, const int64_t* __restrict__ wp, int64_t W,
                            uint64_t* __restrict__ kap, uint64_t* __restrict__ code) {
  const int64_t base = wp[s0];
  for (int64_t q = (int64_t)blockIdx.x * kTB + threadIdx.x; q < W; q += (int64_t)gridDim.x * kTB) {
    const int64_t g = base + q;
    int64_t lo = s0, hi = s1 - 1;  // last s with wp[s] <= g
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (wp[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const int64_t s = lo;
    const int64_t nA = P.seg_start[s + 1] - P.seg_start[s];
    const int64_t r = g - wp[s];
    const int64_t x = r / nA, y = r - x * nA;
    int64_t ix, iy;
    double2 ax, ay;
    pair_member(P, s, x, &ix, &ax);
    pair_member(P, s, y, &iy, &ay);
    kap[q] = key_of_sum(P.scfg + ix * P.n_sites, P.scfg + iy * P.n_sites, 1, P.dims, P.strides, P.n_sites);
    code[q] = ((uint64_t)(s - s0) << (2 * kXBits)) | ((uint64_t)x << kXBits) | (uint64_t)y;
  }
}

__global__ void code_to_seg(const uint64_t* __restrict__ code, int64_t n, uint32_t* __restrict__ seg) {
  for (int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x; t < n; t += (int64_t)gridDim.x * kTB)
    seg[t] = (uint32_t)(code[t] >> (2 * kXBits));
}

__device__ inline void decode(const PairCtx& P, int64_t s0, uint64_t cd, int64_t* s, int64_t* ix, int64_t* iy,
                              double2* v) {
  const int64_t mask = (int64_t(1) << kXBits) - 1;
  *s = s0 + (int64_t)(cd >> (2 * kXBits));
  const int64_t x = (int64_t)(cd >> kXBits) & mask, y = (int64_t)cd & mask;
  double2 ax, ay;
  pair_member(P, *s, x, ix, &ax);
  pair_member(P, *s, y, iy, &ay);
  // A_x conj(A_y)
  *v = make_double2(ax.x * ay.x + ax.y * ay.y, ax.y * ay.x - ax.x * ay.y);
}

// A run starts where the string or kappa changes.
__global__ void run_flags(PairCtx P, int64_t s0, const uint64_t* __restrict__ code, int64_t n,
                          int32_t* __restrict__ flag) {
  for (int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x; t < n; t += (int64_t)gridDim.x * kTB) {
    int f = 1;
    if (t > 0) {
      int64_t s, ix, iy, sp, jx, jy;
      double2 v;
      decode(P, s0, code[t], &s, &ix, &iy, &v);
      decode(P, s0, code[t - 1], &sp, &jx, &jy, &v);
      f = (s != sp) ||
          key_of_sum(P.scfg + ix * P.n_sites, P.scfg + iy * P.n_sites, 1, P.dims, P.strides, P.n_sites) !=
              key_of_sum(P.scfg + jx * P.n_sites, P.scfg + jy * P.n_sites, 1, P.dims, P.strides, P.n_sites);
    }
    flag[t] = f;
  }
}

// |V|^2 of every run, its entries summed in sorted order.
__global__ void run_sums(PairCtx P, int64_t s0, const uint64_t* __restrict__ code, const int64_t* __restrict__ rs,
                         int64_t n_runs, double* __restrict__ out) {
  for (int64_t r = (int64_t)blockIdx.x * kTB + threadIdx.x; r < n_runs; r += (int64_t)gridDim.x * kTB) {
    double2 V = make_double2(0.0, 0.0);
    for (int64_t t = rs[r]; t < rs[r + 1]; ++t) {
      int64_t s, ix, iy;
      double2 v;
      decode(P, s0, code[t], &s, &ix, &iy, &v);
      V.x += v.x;
      V.y += v.y;
    }
    out[r] = V.x * V.x + V.y * V.y;
  }
}

// Fixed-order partial sums: block b adds entries [b * chunk, (b + 1) * chunk)
// by a strided loop and a tree in shared memory.
__global__ void __launch_bounds__(kTB) block_sums(const double* __restrict__ v, int64_t n, int64_t chunk,
                                                  double* __restrict__ part) {
  __shared__ double sh[kTB];
  const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = std::min<int64_t>(n, b0 + chunk);
  double a = 0.0;
  for (int64_t t = b0 + threadIdx.x; t < b1; t += kTB) a += v[t];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int w = kTB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

#define EDLGT_ROCPRIM(call)                                                      \
  do {                                                                           \
    hipError_t _e = (call);                                                      \
    if (_e != hipSuccess) return ::edlgt::hip_status(_e, "rocprim " #call);     \
  } while (0)

// Stable LSD radix sort of (key, value) pairs, in -> out, bits [0, end_bit).
template <class K, class V>
int sort_pairs(const K* kin, K* kout, const V* vin, V* vout, int64_t n, unsigned end_bit, DevBuf& tmp,
               hipStream_t st) {
  size_t bytes = 0;
  EDLGT_ROCPRIM(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)n, 0u, end_bit, st));
  if (bytes > tmp.bytes) EDLGT_TRY(tmp.alloc(bytes));
  EDLGT_ROCPRIM(rocprim::radix_sort_pairs(tmp.p, bytes, kin, kout, vin, vout, (size_t)n, 0u, end_bit, st));
  return EDLGT_OK;
}

// Exclusive scan of 0/1 flags, then the list of head positions with n at the end.
int heads_of(const int32_t* flag, int64_t n, DevBuf& pos, DevBuf& starts, int64_t* count, hipStream_t st) {
  EDLGT_TRY(pos.alloc((n + 1) * sizeof(int64_t)));
  EDLGT_TRY(exclusive_scan_i32_i64(flag, pos.as<int64_t>(), n, st));
  int64_t c = 0;
  EDLGT_HIP(hipMemcpyAsync(&c, pos.as<int64_t>() + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  EDLGT_TRY(starts.alloc((c + 1) * sizeof(int64_t)));
  hipLaunchKernelGGL(scatter_heads, dim3(grid_of(n)), dim3(kTB), 0, st, flag, pos.as<int64_t>(), n,
                     starts.as<int64_t>());
  EDLGT_HIP(hipGetLastError());
  *count = c;
  return EDLGT_OK;
}

// Sum of a device array in a fixed order.
int ordered_sum(const double* v, int64_t n, double* out, hipStream_t st) {
  const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(4096, (n + kTB - 1) / kTB));
  const int64_t chunk = (n + nb - 1) / nb;
  DevBuf part;
  EDLGT_TRY(part.alloc(nb * sizeof(double)));
  hipLaunchKernelGGL(block_sums, dim3((unsigned)nb), dim3(kTB), 0, st, v, n, std::max<int64_t>(chunk, 1),
                     part.as<double>());
  EDLGT_HIP(hipGetLastError());
  std::vector<double> h(nb);
  EDLGT_HIP(hipMemcpyAsync(h.data(), part.p, nb * sizeof(double), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  double s = 0.0;
  for (double x : h) s += x;
  *out = s;
  return EDLGT_OK;
}

}  // namespace

int basis_sre2(const Basis& B, const void* psi_v, double delta, double* sre2, int64_t* support, double* discarded,
               int64_t* n_strings, hipStream_t st) {
  const int64_t N = B.n_states, ns = B.n_sites;
  if (!(delta > 0.0 && delta < 1.0)) return set_error(EDLGT_ERR_INVALID_ARG, "prob_threshold must be 1> delta >0.");
  if (B.key_words != 1)
    return set_error(EDLGT_ERR_INVALID_ARG, "configuration keys above 64 bits are not supported here");
  if (N < 1) return set_error(EDLGT_ERR_INVALID_ARG, "empty sector");
  if (N > 0xffffffffLL) return set_error(EDLGT_ERR_INVALID_ARG, "sectors above 2^32 states are not supported here");
  const double2* psi = static_cast<const double2*>(psi_v);
  // 1. support: probabilities sorted descending (stable), cumulative sum left to right
  DevBuf prob, idx, sprob, sidx, tmp;
  EDLGT_TRY(prob.alloc(N * sizeof(double)));
  EDLGT_TRY(idx.alloc(N * sizeof(uint32_t)));
  EDLGT_TRY(sprob.alloc(N * sizeof(double)));
  EDLGT_TRY(sidx.alloc(N * sizeof(uint32_t)));
  hipLaunchKernelGGL(prob_iota, dim3(grid_of(N)), dim3(kTB), 0, st, psi, N, prob.as<double>(), idx.as<uint32_t>());
  EDLGT_HIP(hipGetLastError());
  {
    size_t bytes = 0;
    EDLGT_ROCPRIM(rocprim::radix_sort_pairs_desc(nullptr, bytes, prob.as<double>(), sprob.as<double>(),
                                                 idx.as<uint32_t>(), sidx.as<uint32_t>(), (size_t)N, 0u, 64u, st));
    EDLGT_TRY(tmp.alloc(bytes));
    EDLGT_ROCPRIM(rocprim::radix_sort_pairs_desc(tmp.p, bytes, prob.as<double>(), sprob.as<double>(),
                                                 idx.as<uint32_t>(), sidx.as<uint32_t>(), (size_t)N, 0u, 64u, st));
  }
  const double target = 1.0 - delta;
  int64_t k_last = N;  // numpy's searchsorted past the end: every state kept
  {
    std::vector<double> h;
    double cum = 0.0;
    const int64_t chunk = 1 << 20;
    for (int64_t b = 0; b < N && k_last == N; b += chunk) {
      const int64_t m = std::min(chunk, N - b);
      h.resize(m);
      EDLGT_HIP(hipMemcpyAsync(h.data(), sprob.as<double>() + b, m * sizeof(double), hipMemcpyDeviceToHost, st));
      EDLGT_HIP(hipStreamSynchronize(st));
      for (int64_t t = 0; t < m; ++t) {
        cum += h[t];
        if (cum >= target) {
          k_last = b + t;
          break;
        }
      }
    }
  }
  const int64_t K = std::min<int64_t>(k_last + 1, N);
  if (K > kSreMaxSupport)
    return set_error(EDLGT_ERR_INVALID_ARG, "support of " + std::to_string(K) + " states (at most " +
                                                std::to_string(kSreMaxSupport) + "): raise prob_threshold");
  // support indices in key order (= index order: the basis is sorted by key)
  DevBuf kidx_s, dummy_s;
  EDLGT_TRY(kidx_s.alloc(K * sizeof(uint32_t)));
  EDLGT_TRY(dummy_s.alloc(K * sizeof(uint32_t)));
  EDLGT_TRY((sort_pairs<uint32_t, uint32_t>(sidx.as<uint32_t>(), kidx_s.as<uint32_t>(), sidx.as<uint32_t>(),
                                           dummy_s.as<uint32_t>(), K, 32u, tmp, st)));
  prob.release();
  idx.release();
  sprob.release();
  DevBuf c, scfg, dims, strides;
  EDLGT_TRY(c.alloc(K * sizeof(double2)));
  EDLGT_TRY(scfg.alloc(K * ns * sizeof(int32_t)));
  hipLaunchKernelGGL(gather_support, dim3(grid_of(K)), dim3(kTB), 0, st, kidx_s.as<uint32_t>(), K, psi,
                     B.configs.as<uint8_t>(), B.cfg_stride, (int)ns, c.as<double2>(), scfg.as<int32_t>());
  EDLGT_HIP(hipGetLastError());
  std::vector<double2> hc(K);
  EDLGT_HIP(hipMemcpyAsync(hc.data(), c.p, K * sizeof(double2), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  double kept = 0.0;
  for (const double2& z : hc) kept += z.x * z.x + z.y * z.y;
  *discarded = std::max(0.0, 1.0 - kept);
  *support = K;
  const double nrm = 1.0 / std::sqrt(kept);
  for (double2& z : hc) z = make_double2(z.x * nrm, z.y * nrm);
  EDLGT_HIP(hipMemcpyAsync(c.p, hc.data(), K * sizeof(double2), hipMemcpyHostToDevice, st));
  EDLGT_TRY(upload(dims, B.loc_dims.data(), ns, st));
  EDLGT_TRY(upload(strides, B.strides.data(), ns, st));
  // 2. X-strings: the K^2 pair keys sorted (stable), grouped by string
  const int64_t NP = K * K;
  DevBuf pk, pay, spk, spay;
  EDLGT_TRY(pk.alloc(NP * sizeof(uint64_t)));
  EDLGT_TRY(pay.alloc(NP * sizeof(uint32_t)));
  EDLGT_TRY(spk.alloc(NP * sizeof(uint64_t)));
  EDLGT_TRY(spay.alloc(NP * sizeof(uint32_t)));
  hipLaunchKernelGGL(string_keys, dim3(grid_of(NP)), dim3(kTB), 0, st, scfg.as<int32_t>(), K, (int)ns,
                     dims.as<int32_t>(), strides.as<uint64_t>(), pk.as<uint64_t>(), pay.as<uint32_t>());
  EDLGT_HIP(hipGetLastError());
  EDLGT_TRY((sort_pairs<uint64_t, uint32_t>(pk.as<uint64_t>(), spk.as<uint64_t>(), pay.as<uint32_t>(),
                                           spay.as<uint32_t>(), NP, 64u, tmp, st)));
  pk.release();
  pay.release();
  DevBuf flag, pos, seg_start;
  EDLGT_TRY(flag.alloc(NP * sizeof(int32_t)));
  hipLaunchKernelGGL(head_flags, dim3(grid_of(NP)), dim3(kTB), 0, st, spk.as<uint64_t>(), NP, flag.as<int32_t>());
  EDLGT_HIP(hipGetLastError());
  int64_t S = 0;
  EDLGT_TRY(heads_of(flag.as<int32_t>(), NP, pos, seg_start, &S, st));
  spk.release();
  *n_strings = S;
  // 3. per string T_p: kappa pairs in batches of whole strings
  DevBuf work, wp;
  EDLGT_TRY(work.alloc(S * sizeof(int64_t)));
  EDLGT_TRY(wp.alloc((S + 1) * sizeof(int64_t)));
  hipLaunchKernelGGL(seg_work, dim3(grid_of(S)), dim3(kTB), 0, st, seg_start.as<int64_t>(), S, work.as<int64_t>());
  EDLGT_HIP(hipGetLastError());
  EDLGT_TRY(exclusive_scan_i64(work.as<int64_t>(), wp.as<int64_t>(), S, st));
  work.release();
  std::vector<int64_t> hwp(S + 1);
  EDLGT_HIP(hipMemcpyAsync(hwp.data(), wp.p, (S + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  int64_t max_w = 0;
  for (int64_t s = 0; s < S; ++s) max_w = std::max(max_w, hwp[s + 1] - hwp[s]);
  const int64_t budget = std::max(kPairBudget, max_w);
  DevBuf kap, code, kap_s, code_s, segk, segk_s, runsq;
  EDLGT_TRY(kap.alloc(budget * sizeof(uint64_t)));
  EDLGT_TRY(code.alloc(budget * sizeof(uint64_t)));
  EDLGT_TRY(kap_s.alloc(budget * sizeof(uint64_t)));
  EDLGT_TRY(code_s.alloc(budget * sizeof(uint64_t)));
  EDLGT_TRY(segk.alloc(budget * sizeof(uint32_t)));
  EDLGT_TRY(segk_s.alloc(budget * sizeof(uint32_t)));
  PairCtx P{seg_start.as<int64_t>(), spay.as<uint32_t>(), c.as<double2>(), scfg.as<int32_t>(),
            dims.as<int32_t>(),      strides.as<uint64_t>(), K, (int)ns};
  double M2 = 0.0;
  for (int64_t s0 = 0; s0 < S;) {
    int64_t s1 = s0 + 1;
    while (s1 < S && hwp[s1 + 1] - hwp[s0] <= budget && s1 - s0 < (int64_t(1) << 30)) ++s1;
    const int64_t W = hwp[s1] - hwp[s0];
    hipLaunchKernelGGL(kappa_pairs, dim3(grid_of(W)), dim3(kTB), 0, st, P, s0, s1, wp.as<int64_t>(), W,
                       kap.as<uint64_t>(), code.as<uint64_t>());
    EDLGT_HIP(hipGetLastError());
    // by kappa, then stably by string: grouped by (string, kappa)
    EDLGT_TRY((sort_pairs<uint64_t, uint64_t>(kap.as<uint64_t>(), kap_s.as<uint64_t>(), code.as<uint64_t>(),
                                             code_s.as<uint64_t>(), W, 64u, tmp, st)));
    hipLaunchKernelGGL(code_to_seg, dim3(grid_of(W)), dim3(kTB), 0, st, code_s.as<uint64_t>(), W,
                       segk.as<uint32_t>());
    EDLGT_HIP(hipGetLastError());
    EDLGT_TRY((sort_pairs<uint32_t, uint64_t>(segk.as<uint32_t>(), segk_s.as<uint32_t>(), code_s.as<uint64_t>(),
                                             code.as<uint64_t>(), W, 32u, tmp, st)));
    DevBuf rflag, rpos, rstart;
    EDLGT_TRY(rflag.alloc(W * sizeof(int32_t)));
    hipLaunchKernelGGL(run_flags, dim3(grid_of(W)), dim3(kTB), 0, st, P, s0, code.as<uint64_t>(), W,
                       rflag.as<int32_t>());
    EDLGT_HIP(hipGetLastError());
    int64_t R = 0;
    EDLGT_TRY(heads_of(rflag.as<int32_t>(), W, rpos, rstart, &R, st));
    EDLGT_TRY(runsq.alloc(R * sizeof(double)));
    hipLaunchKernelGGL(run_sums, dim3(grid_of(R)), dim3(kTB), 0, st, P, s0, code.as<uint64_t>(),
                       rstart.as<int64_t>(), R, runsq.as<double>());
    EDLGT_HIP(hipGetLastError());
    double part = 0.0;
    EDLGT_TRY(ordered_sum(runsq.as<double>(), R, &part, st));
    M2 += part;
    s0 = s1;
  }
  *sre2 = -std::log(std::max(M2, 1e-16));
  return EDLGT_OK;
}

}  // namespace edlgt
