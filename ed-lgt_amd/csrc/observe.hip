// Diagonal local observables on the device.
//
// Every diagonal single-site operator O_i has <psi|O_i|psi> =
// sum_a P_i(a) O_i[a][a], with P_i(a) = sum_r |psi_r|^2 [c_r[i] == a] the
// marginal distribution of site i's local state.  One pass over psi and the
// configuration bytes therefore gives every LocalTerm.get_expval result
// (edlgt/modeling/local_term.py:119-237, exp_val_data qmb_state.py:630-653)
// for all sites and all diagonal operators at once (the reference recomputes
// the triplets of each site and operator per call).
//
// Determinism: |psi_r|^2 is converted to 2^62-scaled fixed point and summed
// with integer atomics in LDS (associative, so the result is independent of
// scheduling); per-block partial histograms are summed in block order.  The
// absolute error is at most N * 2^-62 (< 1e-10 for N < 4.6e8).
#include <algorithm>
#include <mutex>

#include "common.h"

namespace edlgt {
namespace {

constexpr int kMaxBins = 4096;
constexpr double kScale = 4611686018427387904.0;  // 2^62

template <bool kCplx>
__global__ void __launch_bounds__(256) marginal_partial(const double* __restrict__ psi, int64_t n,
                                                       const uint8_t* __restrict__ cfg, int cfg_stride,
                                                       int n_sites, int d_max,
                                                       unsigned long long* __restrict__ partial) {
  __shared__ unsigned long long hist[kMaxBins];
  const int bins = n_sites * d_max;
  for (int i = threadIdx.x; i < bins; i += blockDim.x) hist[i] = 0ull;
  __syncthreads();
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)blockIdx.x * chunk;
  const int64_t r1 = min(n, r0 + chunk);
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    double w;
    if (kCplx) {
      const double re = psi[2 * r], im = psi[2 * r + 1];
      w = re * re + im * im;
    } else {
      w = psi[r] * psi[r];
    }
    const unsigned long long q = (unsigned long long)(w * kScale);
    if (q == 0ull) continue;
    const uint8_t* row = cfg + r * cfg_stride;
    for (int i = 0; i < n_sites; ++i) atomicAdd(&hist[i * d_max + row[i]], q);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < bins; i += blockDim.x)
    partial[(int64_t)blockIdx.x * bins + i] = hist[i];
}

__global__ void marginal_finish(const unsigned long long* __restrict__ partial, int nblk, int bins,
                                double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= bins) return;
  unsigned long long s = 0ull;
  for (int b = 0; b < nblk; ++b) s += partial[(int64_t)b * bins + i];
  out[i] = (double)s / kScale;
}

// Scratch of the allocation-free entry point: per-block partial histograms,
// kScratchWords 64-bit words (4 MiB) per device, held until the process ends.
// Host code never uses the array's name as an address: each device's copy is
// resolved once with hipGetSymbolAddress and kept in g_scratch_dev.
constexpr int64_t kScratchWords = int64_t(1) << 19;
__device__ __align__(256) unsigned long long scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

int marginal_blocks(int64_t n_states, int bins, int64_t max_words) {
  int64_t nblk = std::min<int64_t>(1024, std::max<int64_t>(1, (n_states + 4095) / 4096));
  nblk = std::min<int64_t>(nblk, max_words / bins);
  return (int)nblk;
}

int marginals_launch(const Basis& B, const void* psi, bool cplx, int d_max, double* out_dev,
                     unsigned long long* part, int nblk, hipStream_t st) {
  const int bins = B.n_sites * d_max;
  if (cplx)
    hipLaunchKernelGGL(marginal_partial<true>, dim3(nblk), dim3(256), 0, st, (const double*)psi,
                       B.n_states, B.configs.as<uint8_t>(), B.cfg_stride, B.n_sites, d_max, part);
  else
    hipLaunchKernelGGL(marginal_partial<false>, dim3(nblk), dim3(256), 0, st, (const double*)psi,
                       B.n_states, B.configs.as<uint8_t>(), B.cfg_stride, B.n_sites, d_max, part);
  EDLGT_HIP(hipGetLastError());
  hipLaunchKernelGGL(marginal_finish, dim3((bins + 255) / 256), dim3(256), 0, st, part, nblk, bins,
                     out_dev);
  EDLGT_HIP(hipGetLastError());
  return EDLGT_OK;
}

int check_marginal_args(const Basis& B, int d_max) {
  const int bins = B.n_sites * d_max;
  if (bins > kMaxBins) return set_error(EDLGT_ERR_INVALID_ARG, "n_sites * d_max exceeds 4096");
  for (int s = 0; s < B.n_sites; ++s)
    if (B.loc_dims[s] > d_max) return set_error(EDLGT_ERR_INVALID_ARG, "d_max below a local dimension");
  return EDLGT_OK;
}


// ---- diagonal quarter-turn phases of basis states.
// e_r = (sum_s tab[s * d_max + c_r[s]]) mod 4 and y[v][r] = i^(dir e_r) x[v][r]
// (dir = +1 or -1): maps vectors between a Hamiltonian H and D^H H D, which the
// diagonal phase D = diag(i^e_r) made real (D x for an eigenvector x of the
// real matrix, D^H psi for a state evolved under it).  The per-site table sits
// in LDS; multiplying by a power of i only swaps and negates components, so
// the map is exact.  Real input (kCplxIn false) or complex input; every
// thread reads its own element before writing it, so y may alias x for
// complex input.
template <bool kCplxIn>
__global__ void __launch_bounds__(256) quarter_phase_rows(const uint8_t* __restrict__ cfg, int64_t n,
                                                         int cfg_stride, int n_sites, int d_max,
                                                         const int32_t* __restrict__ tab, int n_vec, int dir,
                                                         const double* x, double2* y) {
  __shared__ int32_t stab[kMaxBins];
  const int bins = n_sites * d_max;
  for (int i = threadIdx.x; i < bins; i += blockDim.x) stab[i] = tab[i] & 3;
  __syncthreads();
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const uint8_t* row = cfg + r * cfg_stride;
    int e = 0;
    for (int s = 0; s < n_sites; ++s) e += stab[s * d_max + row[s]];
    e = (dir < 0 ? 4 - (e & 3) : e) & 3;
    for (int v = 0; v < n_vec; ++v) {
      const int64_t i = (int64_t)v * n + r;
      double2 a;
      if (kCplxIn) a = reinterpret_cast<const double2*>(x)[i];
      else a = make_double2(x[i], 0.0);
      double2 o;
      if (e == 0) o = a;
      else if (e == 1) o = make_double2(-a.y, a.x);
      else if (e == 2) o = make_double2(-a.x, -a.y);
      else o = make_double2(a.y, -a.x);
      y[i] = o;
    }
  }
}
}  // namespace

// out_dev: device double [n_sites * d_max]
int basis_marginals(const Basis& B, const void* psi, bool cplx, int d_max, double* out_dev,
                    hipStream_t st) {
  EDLGT_TRY(check_marginal_args(B, d_max));
  const int bins = B.n_sites * d_max;
  const int nblk = marginal_blocks(B.n_states, bins, int64_t(1024) * bins);
  DevBuf part;
  EDLGT_TRY(part.alloc((size_t)nblk * bins * sizeof(unsigned long long)));
  EDLGT_TRY(marginals_launch(B, psi, cplx, d_max, out_dev, part.as<unsigned long long>(), nblk, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  return EDLGT_OK;
}

// The same histograms with the partial sums in the file-scope scratch array:
// no allocation, no synchronisation (stream-ordered).  Every launch writes all
// nblk * bins partial words before the finish kernel reads them, so nothing
// needs clearing; calls that overlap on one device share the array and race.
// The fixed-point sums are associative, so the result equals basis_marginals'
// bit for bit although the block count may differ.
int basis_marginals_noalloc(const Basis& B, const void* psi, bool cplx, int d_max, double* out_dev,
                            hipStream_t st) {
  EDLGT_TRY(check_marginal_args(B, d_max));
  int dev = 0;
  EDLGT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(EDLGT_ERR_INVALID_ARG, "device ordinal above 63");
  void* part = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
      void* p = nullptr;
      EDLGT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      g_scratch_dev[dev] = p;
    }
    part = g_scratch_dev[dev];
  }
  const int bins = B.n_sites * d_max;
  const int nblk = marginal_blocks(B.n_states, bins, kScratchWords);
  return marginals_launch(B, psi, cplx, d_max, out_dev, static_cast<unsigned long long*>(part), nblk,
                          st);
}

// y (complex, n_vec x N) = i^(dir e) x (x real or complex, n_vec x N), e_r
// from the per-site integer table tab_host [n_sites][d_max] (entries taken
// mod 4).
int basis_quarter_phase(const Basis& B, const int32_t* tab_host, int d_max, int n_vec, int dir, bool cplx_in,
                        const void* x, void* y, hipStream_t st) {
  EDLGT_TRY(check_marginal_args(B, d_max));
  if (n_vec < 0) return set_error(EDLGT_ERR_INVALID_ARG, "negative vector count");
  if (dir != 1 && dir != -1) return set_error(EDLGT_ERR_INVALID_ARG, "direction must be +1 or -1");
  if (B.n_states == 0 || n_vec == 0) return EDLGT_OK;
  const int bins = B.n_sites * d_max;
  DevBuf tab;
  EDLGT_TRY(tab.alloc((size_t)bins * sizeof(int32_t)));
  EDLGT_HIP(hipMemcpyAsync(tab.p, tab_host, (size_t)bins * sizeof(int32_t), hipMemcpyHostToDevice, st));
  const unsigned g = grid_for(B.n_states, 256, 16384);
  if (cplx_in)
    hipLaunchKernelGGL(quarter_phase_rows<true>, dim3(g), dim3(256), 0, st, B.configs.as<uint8_t>(), B.n_states,
                       B.cfg_stride, B.n_sites, d_max, tab.as<int32_t>(), n_vec, dir, static_cast<const double*>(x),
                       static_cast<double2*>(y));
  else
    hipLaunchKernelGGL(quarter_phase_rows<false>, dim3(g), dim3(256), 0, st, B.configs.as<uint8_t>(), B.n_states,
                       B.cfg_stride, B.n_sites, d_max, tab.as<int32_t>(), n_vec, dir, static_cast<const double*>(x),
                       static_cast<double2*>(y));
  EDLGT_HIP(hipGetLastError());
  EDLGT_HIP(hipStreamSynchronize(st));  // tab is freed on return
  return EDLGT_OK;
}

namespace {

// rho[r] = sum_b |sum_{t in block b} coef[t] E[rows[t]][r]|^2, one thread per
// basis state r (coalesced reads of the eigenvector rows), blocks and their
// members in the order given (deterministic).
__global__ void __launch_bounds__(256) ensemble_density_rows(const double2* __restrict__ E, int64_t n,
                                                            const int64_t* __restrict__ rows,
                                                            const double2* __restrict__ coef,
                                                            const int64_t* __restrict__ bstart, int64_t n_blocks,
                                                            double* __restrict__ rho) {
  for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256) {
    double acc = 0.0;
    for (int64_t b = 0; b < n_blocks; ++b) {
      double sx = 0.0, sy = 0.0;
      for (int64_t t = bstart[b]; t < bstart[b + 1]; ++t) {
        const double2 e = E[rows[t] * n + r], c = coef[t];
        sx += c.x * e.x - c.y * e.y;
        sy += c.x * e.y + c.y * e.x;
      }
      acc += sx * sx + sy * sy;
    }
    rho[r] = acc;
  }
}

}  // namespace

int ensemble_density(const void* E, int64_t n, int64_t n_rows_E, int64_t k, const int64_t* rows_host,
                     const double* coef_host, const int64_t* bstart_host, int64_t n_blocks, double* rho_dev,
                     hipStream_t st) {
  if (n < 0 || k < 0 || n_blocks < 0) return set_error(EDLGT_ERR_INVALID_ARG, "negative size");
  if (n == 0) return EDLGT_OK;
  if (!E || !rho_dev || (k && (!rows_host || !coef_host)) || !bstart_host)
    return set_error(EDLGT_ERR_INVALID_ARG, "null argument");
  if (bstart_host[0] != 0 || bstart_host[n_blocks] != k)
    return set_error(EDLGT_ERR_INVALID_ARG, "block starts must run from 0 to the member count");
  for (int64_t b = 0; b < n_blocks; ++b)
    if (bstart_host[b + 1] < bstart_host[b]) return set_error(EDLGT_ERR_INVALID_ARG, "block starts must not decrease");
  for (int64_t t = 0; t < k; ++t)
    if (rows_host[t] < 0 || rows_host[t] >= n_rows_E)
      return set_error(EDLGT_ERR_INVALID_ARG, "member row " + std::to_string(rows_host[t]) + " outside the matrix");
  DevBuf rows, coef, bs;
  EDLGT_TRY(upload(rows, rows_host, (size_t)k, st));
  EDLGT_TRY(upload(coef, reinterpret_cast<const double2*>(coef_host), (size_t)k, st));
  EDLGT_TRY(upload(bs, bstart_host, (size_t)n_blocks + 1, st));
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 65535);
  hipLaunchKernelGGL(ensemble_density_rows, dim3(g), dim3(256), 0, st, static_cast<const double2*>(E), n,
                     rows.as<int64_t>(), coef.as<double2>(), bs.as<int64_t>(), n_blocks, rho_dev);
  EDLGT_HIP(hipGetLastError());
  EDLGT_HIP(hipStreamSynchronize(st));  // the member lists are freed on return
  return EDLGT_OK;
}

}  // namespace edlgt
