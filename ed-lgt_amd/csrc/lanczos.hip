// Block thick-restart Lanczos for the lowest eigenpairs of a Hermitian CSR
// matrix (replaces scipy eigsh(H, k, which="SA"), edlgt/modeling/
// qmb_hamiltonian.py:260-332).
//
// Algorithm (Wu & Simon thick restart, block form):
//   V = [V_0 .. V_{nv-1}] orthonormal (column-major, N x (M + b) on the
//   device); the active block A = V[:, nv-b:nv] is multiplied by H with one
//   SpMM (the matrix is streamed once for all b vectors), orthogonalised
//   against the whole basis with two classical Gram-Schmidt passes
//   (G = V^H W, W -= V G), and factored W = Q R by Cholesky QR twice.  The
//   projected matrix T = V^H H V (Hermitian, host) receives G and R.  When
//   the basis is full, T is diagonalised (Householder + implicit QL), residual
//   norms ||R S_i|| decide convergence, and the l lowest Ritz vectors are
//   kept (V <- V S[:, :l]) together with the residual block Q and the
//   coupling R S.  Rank loss of W (invariant subspace) is repaired column by
//   column with random vectors that carry no coupling.
// Device kernels: tall-skinny Gram products V^H W with a deterministic
// two-stage reduction, tall-skinny updates W = beta W + V C, and the CSR SpMM.
// Mixed precision: the SpMM reads FP32 values and an FP32 copy of the active
// block; every dot, the recurrence, the reorthogonalisation and T stay FP64.
// The FP32 phase is followed by an FP64 phase started from its Ritz vectors.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "common.h"
#include "gram.h"
#include "kernels.h"

namespace edlgt {
namespace {

using namespace gramk;
using cd = std::complex<double>;

constexpr double kDgks = 0.5;  // squared-norm ratio above which one Gram-Schmidt pass suffices
constexpr double kSkipLoss = 1e-14;  // measured loss of orthogonality left uncorrected
// Mixed precision, one eigenpair: relative residual of the FP32-product
// Lanczos phase, the relative CG residual of each Newton correction, and the
// CG iterations allowed per correction (EDLGT_MIXED_TOL32 / EDLGT_MIXED_ETA
// override the first two for tuning runs)
constexpr double kMixedTol32 = 1e-3;
constexpr double kMixedEta = 1e-4;
constexpr int kCgMaxIter = 600;
// Krylov exponential: a step's basis stops growing once its a-posteriori
// error estimate (relative to |psi|) is below kKrylovTol, tested from
// kKrylovMin vectors on (the requested m is the maximum)
constexpr double kKrylovTol = 1e-13;
constexpr int kKrylovMin = 4;
double env_double(const char* name, double dflt) {
  const char* e = std::getenv(name);
  return e && *e ? std::atof(e) : dflt;
}

// Per-thread Gram accumulator of the NW values a thread writes for one row
// (g[i][c] += conj(y_i) y_c), reduced over the block at the end (fixed wave
// butterfly, waves in order) into partial[blk][i][c].
template <class S, int NW>
struct RowGram {
  S g[NW][NW];
  __device__ void clear() {
#pragma unroll
    for (int i = 0; i < NW; ++i)
#pragma unroll
      for (int c = 0; c < NW; ++c) g[i][c] = zero<S>();
  }
  __device__ void add_row(const S (&y)[NW]) {
#pragma unroll
    for (int i = 0; i < NW; ++i)
#pragma unroll
      for (int c = 0; c < NW; ++c) g[i][c] = add(g[i][c], cj_mul(y[i], y[c]));
  }
  __device__ void flush(double2* __restrict__ partial) {
    __shared__ double2 red[kTB / 64][NW * NW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NW; ++i)
#pragma unroll
      for (int c = 0; c < NW; ++c) {
        const S t = wave_sum(g[i][c]);
        if (lane == 0) red[wave][i * NW + c] = to_c(t);
      }
    __syncthreads();
    for (int t = threadIdx.x; t < NW * NW; t += kTB) {
      double2 s = red[0][t];
#pragma unroll
      for (int q = 1; q < kTB / 64; ++q) s = add(s, red[q][t]);
      partial[(int64_t)blockIdx.x * NW * NW + t] = s;
    }
  }
};

// Rows per thread per loop iteration of the streaming updates: independent
// rows whose loads are all issued before any is used, so that a wave keeps
// several memory requests in flight (one row per iteration ran the QED 4x3
// local update at 3.2 TB/s).  A thread still visits its rows in the order
// r, r + stride, r + 2 stride, ..., so its Gram sums are added in the same
// order as with one row per iteration: results are unchanged bit for bit.
template <int NW> constexpr int rows_unroll() { return NW <= 2 ? 4 : NW <= 4 ? 2 : 1; }

// W[:, c] = beta * W[:, c] + sum_j V_j C[j][c]    (C: nv x nw, row-major, complex)
// With kGram the Gram matrix of the new W columns is accumulated on the way
// (partial[blk][i][c], NW x NW per block).
template <class S, int NW, bool kGram>
__global__ void __launch_bounds__(kTB) tall_update(const S* __restrict__ V, int64_t ldv, int nv,
                                                   const double2* __restrict__ C, int ldc,
                                                   S* __restrict__ W, int64_t ldw, int nw,
                                                   int64_t n, int beta, double2* __restrict__ partial) {
  constexpr int RU = rows_unroll<NW>();
  extern __shared__ double2 cs[];
  for (int i = threadIdx.x; i < nv * nw; i += blockDim.x) cs[i] = C[(i / nw) * ldc + (i % nw)];
  __syncthreads();
  RowGram<S, kGram ? NW : 1> rg;
  if (kGram) rg.clear();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r0 < n; r0 += stride * RU) {
    S acc[RU][NW];
#pragma unroll
    for (int q = 0; q < RU; ++q) {
      const int64_t r = r0 + q * stride;
#pragma unroll
      for (int c = 0; c < NW; ++c) acc[q][c] = (beta && c < nw && r < n) ? W[c * ldw + r] : zero<S>();
    }
    for (int j = 0; j < nv; ++j) {
      S v[RU];
#pragma unroll
      for (int q = 0; q < RU; ++q) {
        const int64_t r = r0 + q * stride;
        v[q] = r < n ? V[(int64_t)j * ldv + r] : zero<S>();
      }
#pragma unroll
      for (int q = 0; q < RU; ++q)
#pragma unroll
        for (int c = 0; c < NW; ++c)
          if (c < nw) acc[q][c] = add(acc[q][c], mul_s(v[q], cs[j * nw + c]));
    }
#pragma unroll
    for (int q = 0; q < RU; ++q) {
      const int64_t r = r0 + q * stride;
      if (r >= n) continue;
#pragma unroll
      for (int c = 0; c < NW; ++c)
        if (c < nw) W[c * ldw + r] = acc[q][c];
      if constexpr (kGram) rg.add_row(acc[q]);
    }
  }
  if constexpr (kGram) rg.flush(partial);
}

// In place X[:, c] = sum_j X[:, j] C[j][c] for c, j < nw (each thread reads
// its row's nw values before writing them); optionally accumulates the Gram
// matrix of the new columns like tall_update.
template <class S, int NW, bool kGram>
__global__ void __launch_bounds__(kTB) row_transform(S* X, int64_t ldx, int nw, const double2* __restrict__ C,
                                                     int64_t n, double2* __restrict__ partial) {
  constexpr int RU = rows_unroll<NW>();
  __shared__ double2 cs[NW * NW];
  for (int i = threadIdx.x; i < nw * nw; i += blockDim.x) cs[i] = C[i];
  __syncthreads();
  RowGram<S, kGram ? NW : 1> rg;
  if (kGram) rg.clear();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r0 < n; r0 += stride * RU) {
    S x[RU][NW];
#pragma unroll
    for (int q = 0; q < RU; ++q) {
      const int64_t r = r0 + q * stride;
#pragma unroll
      for (int j = 0; j < NW; ++j) x[q][j] = (j < nw && r < n) ? X[j * ldx + r] : zero<S>();
    }
#pragma unroll
    for (int q = 0; q < RU; ++q) {
      const int64_t r = r0 + q * stride;
      if (r >= n) continue;
      S y[NW];
#pragma unroll
      for (int c = 0; c < NW; ++c) {
        y[c] = zero<S>();
#pragma unroll
        for (int j = 0; j < NW; ++j)
          if (j < nw && c < nw) y[c] = add(y[c], mul_s(x[q][j], cs[j * nw + c]));
      }
#pragma unroll
      for (int c = 0; c < NW; ++c)
        if (c < nw) X[c * ldx + r] = y[c];
      if constexpr (kGram) rg.add_row(y);
    }
  }
  if constexpr (kGram) rg.flush(partial);
}

__device__ inline double2 rscale(double c, double2 v) { return make_double2(c * v.x, c * v.y); }

// One three-term Krylov update in place: w = c2 w + c1 b + c0 a (real
// coefficients; a may alias b), accumulating |w|^2 per block like
// tall_update.  Reads three vectors and writes one: the recurrence, the new
// norm and the next basis vector in a single pass.
template <class S>
__global__ void __launch_bounds__(kTB) three_term(const S* a, const S* b, S* w, double c0, double c1, double c2,
                                                  int64_t n, double2* __restrict__ partial) {
  constexpr int RU = 4;
  RowGram<S, 1> rg;
  rg.clear();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r0 < n; r0 += stride * RU) {
    S wa[RU], wb[RU], ww[RU];
#pragma unroll
    for (int q = 0; q < RU; ++q) {
      const int64_t r = r0 + q * stride;
      if (r < n) {
        ww[q] = w[r];
        wb[q] = b[r];
        wa[q] = a[r];
      }
    }
#pragma unroll
    for (int q = 0; q < RU; ++q) {
      const int64_t r = r0 + q * stride;
      if (r >= n) continue;
      S y[1];
      y[0] = add(add(rscale(c2, ww[q]), rscale(c1, wb[q])), rscale(c0, wa[q]));
      w[r] = y[0];
      rg.add_row(y);
    }
  }
  rg.flush(partial);
}

// ---------------------------------------------------------------- CG refinement
// Conjugate gradients on the projected correction equation of one eigenpair
// (P (H32 - theta) P) delta = -r, P = I - x x^H, with the FP32 matrix H32 (the
// mixed-precision products).  The CG vectors (r, p, q, delta, g and a copy of
// x) are stored in FP32 (L = float2 / float) and combined in FP64 registers:
// the correction only needs a relative accuracy eta >= 1e-6, and halving
// their bytes halves the streaming kernels' time.  Every scalar of an
// iteration (beta, gamma, alpha) is reduced on the device from the previous
// kernel's per-block partial sums, each block summing the kCgBlocks partials
// in the same fixed order, so an iteration is three streaming kernels and one
// SpMV with no host synchronisation; the host reads |r|^2 every few
// iterations.
constexpr int kCgBlocks = 1024;

__device__ inline double2 sum_partials(const double2* __restrict__ part) {
  __shared__ double2 red[kTB];
  double2 v = make_double2(0.0, 0.0);
  for (int i = threadIdx.x; i < kCgBlocks; i += kTB) v = add(v, part[i]);
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = kTB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = add(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const double2 out = red[0];
  __syncthreads();
  return out;
}

__device__ inline void store_block_sum(double2 v, double2* __restrict__ part) {
  __shared__ double2 red[kTB / 64];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 s = red[0];
#pragma unroll
    for (int q = 1; q < kTB / 64; ++q) s = add(s, red[q]);
    part[blockIdx.x] = s;
  }
}

__device__ inline float2 to_lo(double2 v) { return make_float2((float)v.x, (float)v.y); }
__device__ inline float to_lo(double v) { return (float)v; }
__device__ inline double2 to_hi(float2 v) { return make_double2((double)v.x, (double)v.y); }
__device__ inline double to_hi(float v) { return (double)v; }
__device__ inline double2 axpy(double a, double2 x, double2 y) { return make_double2(y.x + a * x.x, y.y + a * x.y); }
__device__ inline double axpy(double a, double x, double y) { return y + a * x; }
__device__ inline double2 cmul_s(double2 c, double2 x) { return mul(c, x); }  // complex scalar times x
__device__ inline double cmul_s(double2 c, double x) { return c.x * x; }
__device__ inline double nrm2(double2 v) { return v.x * v.x + v.y * v.y; }
__device__ inline double nrm2(double v) { return v * v; }
__device__ inline double2 sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double sub(double a, double b) { return a - b; }
__device__ inline double2 neg(double2 a) { return make_double2(-a.x, -a.y); }
__device__ inline double neg(double a) { return -a; }

// partials of |hx - theta x|^2 (the FP64 residual norm, not stored)
template <class S>
__global__ void __launch_bounds__(kTB) resid_rr(const S* __restrict__ hx, const S* __restrict__ x, double theta,
                                                int64_t n, double2* __restrict__ part) {
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTB)
    acc += nrm2(axpy(-theta, x[i], hx[i]));
  store_block_sum(make_double2(acc, 0.0), part);
}

// Start of one correction: g = y - theta x (y = H32 x_lo, the product the CG
// operator projects with), r = -(hx - theta x) (the FP64 residual), delta = 0,
// and the partials of |r|^2.
template <class S, class L>
__global__ void __launch_bounds__(kTB) cg_init(const S* __restrict__ y, const S* __restrict__ hx,
                                               const S* __restrict__ x, double theta, int64_t n, L* __restrict__ g,
                                               L* __restrict__ r, L* __restrict__ delta,
                                               double2* __restrict__ part_rr) {
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTB) {
    const S xv = x[i];
    g[i] = to_lo(axpy(-theta, xv, y[i]));
    const L rv = to_lo(neg(axpy(-theta, xv, hx[i])));
    r[i] = rv;
    delta[i] = to_lo(zero<S>());
    acc += nrm2(to_hi(rv));
  }
  store_block_sum(make_double2(acc, 0.0), part_rr);
}

// p = r + beta p (beta = |r|^2 / |r_prev|^2, 0 on the first iteration) and the
// partials of gamma = g^H p
template <class L>
__global__ void __launch_bounds__(kTB) cg_dir(const L* __restrict__ r, L* __restrict__ p, const L* __restrict__ g,
                                              const double2* __restrict__ part_rr,
                                              const double2* __restrict__ part_rr_prev, int first, int64_t n,
                                              double2* __restrict__ part_g) {
  double beta = 0.0;
  if (!first) {
    const double rr = sum_partials(part_rr).x, rp = sum_partials(part_rr_prev).x;
    beta = rp > 0.0 ? rr / rp : 0.0;
  }
  double2 acc = make_double2(0.0, 0.0);
  for (int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTB) {
    const L v = first ? r[i] : to_lo(axpy(beta, to_hi(p[i]), to_hi(r[i])));
    p[i] = v;
    acc = add(acc, to_c(cj_mul(to_hi(g[i]), to_hi(v))));
  }
  store_block_sum(acc, part_g);
}

// q = y - theta p - gamma x (y = H32 p) and the partials of p^H q
template <class S, class L>
__global__ void __launch_bounds__(kTB) cg_apply(const L* __restrict__ p, const S* __restrict__ y,
                                                const L* __restrict__ x, L* __restrict__ q,
                                                const double2* __restrict__ part_g, double theta, int64_t n,
                                                double2* __restrict__ part_pq) {
  const double2 gamma = sum_partials(part_g);
  double2 acc = make_double2(0.0, 0.0);
  for (int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTB) {
    const S pv = to_hi(p[i]);
    const S qv = sub(axpy(-theta, pv, y[i]), cmul_s(gamma, to_hi(x[i])));
    q[i] = to_lo(qv);
    acc = add(acc, to_c(cj_mul(pv, qv)));
  }
  store_block_sum(acc, part_pq);
}

// alpha = |r|^2 / p^H q; delta += alpha p; r -= alpha q; partials of the new |r|^2
// (alpha = 0 when p^H q is not positive: the host sees |r| stall and stops)
template <class L>
__global__ void __launch_bounds__(kTB) cg_step(L* __restrict__ delta, L* __restrict__ r, const L* __restrict__ p,
                                               const L* __restrict__ q, const double2* __restrict__ part_rr,
                                               const double2* __restrict__ part_pq, int64_t n,
                                               double2* __restrict__ part_rr_out) {
  const double rr = sum_partials(part_rr).x, pq = sum_partials(part_pq).x;
  const double alpha = pq > 0.0 ? rr / pq : 0.0;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTB) {
    delta[i] = to_lo(axpy(alpha, to_hi(p[i]), to_hi(delta[i])));
    const L rv = to_lo(axpy(-alpha, to_hi(q[i]), to_hi(r[i])));
    r[i] = rv;
    acc += nrm2(to_hi(rv));
  }
  store_block_sum(make_double2(acc, 0.0), part_rr_out);
}

// x += delta and the partials of the new |x|^2
template <class S, class L>
__global__ void __launch_bounds__(kTB) cg_accept(S* __restrict__ x, const L* __restrict__ delta, int64_t n,
                                                 double2* __restrict__ part) {
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTB) {
    const S v = add(x[i], to_hi(delta[i]));
    x[i] = v;
    acc += nrm2(v);
  }
  store_block_sum(make_double2(acc, 0.0), part);
}

// fills x[i] with deterministic pseudo-random values in [-0.5, 0.5)
__device__ inline double hash_unit(uint64_t x) {
  x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
  return (double)(x >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}
template <class S> __device__ inline S rand_val(uint64_t seed, int64_t i);
template <> __device__ inline double rand_val<double>(uint64_t seed, int64_t i) {
  return hash_unit(seed * 0x9e3779b97f4a7c15ull + (uint64_t)i);
}
template <> __device__ inline double2 rand_val<double2>(uint64_t seed, int64_t i) {
  return make_double2(hash_unit(seed * 0x9e3779b97f4a7c15ull + 2 * (uint64_t)i),
                      hash_unit(seed * 0x9e3779b97f4a7c15ull + 2 * (uint64_t)i + 1));
}
template <class S>
__global__ void fill_random(S* __restrict__ x, int64_t n, uint64_t seed) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = rand_val<S>(seed, i);
}

// ---------------------------------------------------------------- host side
struct Dense {  // small dense complex matrix, row-major
  int r = 0, c = 0;
  std::vector<cd> a;
  Dense() = default;
  Dense(int rr, int cc) : r(rr), c(cc), a((size_t)rr * cc, cd(0, 0)) {}
  cd& operator()(int i, int j) { return a[(size_t)i * c + j]; }
  cd operator()(int i, int j) const { return a[(size_t)i * c + j]; }
};

// Eigenvalues (ascending) and eigenvectors (the columns of S) of a small
// Hermitian matrix in O(n^3): Householder reduction to a Hermitian
// tridiagonal matrix, a diagonal phase change that makes it real symmetric,
// the implicit QL iteration with Wilkinson shifts on that (Golub & Van Loan,
// Matrix Computations, 4th ed., 8.3), and the back-transformation.  The
// projected matrices of the solvers have up to a few hundred rows and are
// factored at every block step: at n = 64 / 128 this takes 1.4 / 10.5 ms on
// one host core, where cyclic Jacobi took 9.6 / 112 ms.
void hermitian_eig(const Dense& A0, std::vector<double>& w, Dense& S) {
  const int n = A0.r;
  Dense A = A0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {  // enforce exact Hermitian symmetry
      const cd m = 0.5 * (A(i, j) + std::conj(A(j, i)));
      A(i, j) = m;
      A(j, i) = std::conj(m);
    }
  for (int i = 0; i < n; ++i) A(i, i) = cd(A(i, i).real(), 0.0);
  Dense Q(n, n);
  for (int i = 0; i < n; ++i) Q(i, i) = 1.0;
  std::vector<cd> v(n), p(n), q(n);
  // Householder: H_k = I - 2 u u^H (|u| = 1) maps A[k+1:, k] to alpha e_1;
  // A <- H_k A H_k on the trailing block, Q <- Q H_k.
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    double xn2 = 0.0;
    for (int i = 0; i < m; ++i) xn2 += std::norm(A(k + 1 + i, k));
    if (xn2 == 0.0) continue;
    const double xn = std::sqrt(xn2);
    const cd x0 = A(k + 1, k);
    const double ax0 = std::abs(x0);
    const cd alpha = -(ax0 > 0.0 ? x0 / ax0 : cd(1.0, 0.0)) * xn;
    for (int i = 0; i < m; ++i) v[i] = A(k + 1 + i, k);
    v[0] -= alpha;
    double vn2 = 0.0;
    for (int i = 0; i < m; ++i) vn2 += std::norm(v[i]);
    if (vn2 == 0.0) continue;
    const double vs = 1.0 / std::sqrt(vn2);
    for (int i = 0; i < m; ++i) v[i] *= vs;
    // (the inner loops spell out complex products in real arithmetic: the
    // library's complex multiply guards infinities and is several times
    // slower)
    double K = 0.0;  // u^H A22 u (real)
    for (int i = 0; i < m; ++i) {
      const cd* ar = &A.a[(size_t)(k + 1 + i) * n + k + 1];
      double zr = 0.0, zi = 0.0;
      for (int j = 0; j < m; ++j) {
        zr += ar[j].real() * v[j].real() - ar[j].imag() * v[j].imag();
        zi += ar[j].real() * v[j].imag() + ar[j].imag() * v[j].real();
      }
      p[i] = cd(zr, zi);
      K += v[i].real() * zr + v[i].imag() * zi;
    }
    for (int i = 0; i < m; ++i) q[i] = cd(2.0 * (p[i].real() - K * v[i].real()), 2.0 * (p[i].imag() - K * v[i].imag()));
    for (int i = 0; i < m; ++i) {
      cd* ar = &A.a[(size_t)(k + 1 + i) * n + k + 1];
      const double vr = v[i].real(), vi = v[i].imag(), qr = q[i].real(), qi = q[i].imag();
      for (int j = 0; j < m; ++j) {
        // v_i conj(q_j) + q_i conj(v_j)
        const double tr = vr * q[j].real() + vi * q[j].imag() + qr * v[j].real() + qi * v[j].imag();
        const double ti = vi * q[j].real() - vr * q[j].imag() + qi * v[j].real() - qr * v[j].imag();
        ar[j] = cd(ar[j].real() - tr, ar[j].imag() - ti);
      }
    }
    for (int i = 0; i < m; ++i) {
      A(k + 1 + i, k) = i == 0 ? alpha : cd(0.0, 0.0);
      A(k, k + 1 + i) = std::conj(A(k + 1 + i, k));
    }
    for (int r = 0; r < n; ++r) {
      cd* qr = &Q.a[(size_t)r * n + k + 1];
      double zr = 0.0, zi = 0.0;
      for (int j = 0; j < m; ++j) {
        zr += qr[j].real() * v[j].real() - qr[j].imag() * v[j].imag();
        zi += qr[j].real() * v[j].imag() + qr[j].imag() * v[j].real();
      }
      zr *= 2.0;
      zi *= 2.0;
      for (int j = 0; j < m; ++j) {  // -= z conj(v_j)
        qr[j] = cd(qr[j].real() - (zr * v[j].real() + zi * v[j].imag()),
                   qr[j].imag() - (zi * v[j].real() - zr * v[j].imag()));
      }
    }
  }
  // T = D T' D^H with T' real: diagonal d, off-diagonal e[i] = |T(i+1, i)|
  std::vector<double> d(n), e(n, 0.0);
  std::vector<cd> ph(n, cd(1.0, 0.0));
  for (int i = 0; i < n; ++i) d[i] = A(i, i).real();
  for (int i = 0; i + 1 < n; ++i) {
    const cd o = A(i + 1, i);
    e[i] = std::abs(o);
    ph[i + 1] = e[i] > 0.0 ? ph[i] * (o / e[i]) : ph[i];
  }
  // implicit QL with Wilkinson shifts; Zt (the transpose of the eigenvector
  // matrix, so that a rotation of two columns touches two contiguous rows)
  // accumulates the rotations
  std::vector<double> Zt((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Zt[(size_t)i * n + i] = 1.0;
  for (int l = 0; l < n; ++l) {
    for (int iter = 0; iter < 64; ++iter) {
      int m = l;
      for (; m + 1 < n; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 2.2204460492503131e-16 * dd) break;
      }
      if (m == l) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
      double s = 1.0, c = 1.0, pp = 0.0;
      bool deflated = false;
      for (int i = m - 1; i >= l; --i) {
        double f = s * e[i];
        const double b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {  // underflow: split the matrix here and start again
          d[i + 1] -= pp;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - pp;
        r = (d[i] - g) * s + 2.0 * c * b;
        pp = s * r;
        d[i + 1] = g + pp;
        g = c * r - b;
        double* za = &Zt[(size_t)i * n];
        double* zb = &Zt[(size_t)(i + 1) * n];
        for (int k = 0; k < n; ++k) {
          const double t = zb[k];
          zb[k] = s * za[k] + c * t;
          za[k] = c * za[k] - s * t;
        }
      }
      if (deflated) continue;
      d[l] -= pp;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return d[x] < d[y]; });
  // eigenvectors of A: Q D Z, columns in ascending order
  S = Dense(n, n);
  w.assign(n, 0.0);
  std::vector<double> yr(n), yi(n);
  for (int kk = 0; kk < n; ++kk) {
    const int k = ord[kk];
    w[kk] = d[k];
    const double* zk = &Zt[(size_t)k * n];
    for (int i = 0; i < n; ++i) {
      yr[i] = ph[i].real() * zk[i];
      yi[i] = ph[i].imag() * zk[i];
    }
    for (int r = 0; r < n; ++r) {
      const cd* qr = &Q.a[(size_t)r * n];
      double zr = 0.0, zi = 0.0;
      for (int i = 0; i < n; ++i) {
        zr += qr[i].real() * yr[i] - qr[i].imag() * yi[i];
        zi += qr[i].real() * yi[i] + qr[i].imag() * yr[i];
      }
      S(r, kk) = cd(zr, zi);
    }
  }
}

// Cholesky of a small Hermitian positive matrix: G = R^H R, R upper.
// Returns false if a pivot falls below floor2[c].
bool cholesky_upper(const Dense& G, Dense& R, const std::vector<double>& floor2) {
  const int n = G.r;
  R = Dense(n, n);
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i <= j; ++i) {
      cd s = G(i, j);
      for (int k = 0; k < i; ++k) s -= std::conj(R(k, i)) * R(k, j);
      if (i == j) {
        const double d = s.real();
        if (!(d > floor2[j])) return false;
        R(j, j) = std::sqrt(d);
      } else {
        R(i, j) = s / R(i, i).real();
      }
    }
  }
  return true;
}

Dense upper_inverse(const Dense& R) {
  const int n = R.r;
  Dense X(n, n);
  for (int j = 0; j < n; ++j) {
    X(j, j) = 1.0 / R(j, j);
    for (int i = j - 1; i >= 0; --i) {
      cd s = 0.0;
      for (int k = i + 1; k <= j; ++k) s += R(i, k) * X(k, j);
      X(i, j) = -s / R(i, i);
    }
  }
  return X;
}

Dense matmul(const Dense& A, const Dense& B) {
  Dense C(A.r, B.c);
  for (int i = 0; i < A.r; ++i)
    for (int k = 0; k < A.c; ++k) {
      const cd a = A(i, k);
      if (a == cd(0, 0)) continue;
      for (int j = 0; j < B.c; ++j) C(i, j) += a * B(k, j);
    }
  return C;
}

// Page-locked, coherent host memory (hipHostMalloc) that the Gram reduction
// kernels write directly (device address `dev`): the host reads a few KB of
// Gram entries after a stream synchronisation with no copy in between
// (device stores to coherent host memory bypass the GPU caches).
struct HostPinned {
  void* p = nullptr;
  void* dev = nullptr;
  size_t bytes = 0;
  HostPinned() = default;
  HostPinned(const HostPinned&) = delete;
  HostPinned& operator=(const HostPinned&) = delete;
  ~HostPinned() {
    if (p) (void)hipHostFree(p);
  }
  int alloc(size_t n) {
    if (p) (void)hipHostFree(p);
    p = dev = nullptr;
    bytes = 0;
    if (hipHostMalloc(&p, n, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) {
      p = nullptr;
      return set_error(EDLGT_ERR_OUT_OF_MEMORY, "hipHostMalloc of " + std::to_string(n) + " bytes failed");
    }
    if (hipHostGetDevicePointer(&dev, p, 0) != hipSuccess || !dev) {
      (void)hipHostFree(p);
      p = dev = nullptr;
      return set_error(EDLGT_ERR_HIP, "hipHostGetDevicePointer failed on pinned Gram memory");
    }
    bytes = n;
    return EDLGT_OK;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
  template <class T> T* dev_as() const { return static_cast<T*>(dev); }
};

template <class S>
struct Solver {
  Csr& h;
  hipStream_t st;
  int precision;
  int64_t N;
  int b;       // block size
  int M;       // basis capacity in vectors (multiple of b)
  DevBuf V, W, Tmp, part, cdev, shadow, pack;
  HostPinned hgram;  // Gram entries, written by gram_finish
  int64_t mv64 = 0, mv32 = 0;
  int64_t ortho_single = 0, ortho_double = 0;  // orthogonalisations with one / two passes
  int64_t ortho_skipped = 0;  // full passes whose update was skipped (loss below kSkipLoss)
  uint64_t seed = 12345;
  bool need_tmp = true;      // restart / Ritz buffer (not needed by Krylov stepping)
  bool want_vectors = true;  // form the Ritz vectors when a run finishes
  bool need_w = true;        // spare block (Krylov stepping, true residuals)
  // Orthogonalise each new block against the last two blocks only (after the
  // first step of a start or restart).  Cheap and accurate while no Ritz pair
  // has converged far (residuals well above sqrt(unit roundoff) of the
  // products): used for the loose FP32 phase of the mixed mode, whose result
  // the FP64 refinement checks.
  bool local_only = false;

  Solver(Csr& hh, hipStream_t s, int prec) : h(hh), st(s), precision(prec), N(hh.n_rows) {}

  S* col(int j) { return V.as<S>() + (int64_t)j * N; }

  int setup() {
    EDLGT_TRY(V.alloc((size_t)(M + b) * N * sizeof(S)));
    if (need_w) EDLGT_TRY(W.alloc((size_t)b * N * sizeof(S)));
    if (need_tmp) EDLGT_TRY(Tmp.alloc((size_t)std::max(M + b, 8) * N * sizeof(S)));
    EDLGT_TRY(part.alloc((size_t)kGramBlocks * kGramCols * 8 * sizeof(double2)));
    EDLGT_TRY(hgram.alloc((size_t)(M + b) * std::max(M, 8) * sizeof(double2)));
    EDLGT_TRY(cdev.alloc((size_t)(M + b) * std::max(M, 8) * sizeof(double2)));
    if (precision == EDLGT_PREC_MIXED) {
      EDLGT_TRY(csr_make_low_precision(h, st));
      EDLGT_TRY(shadow.alloc((size_t)b * N * sizeof(S) / 2));
    }
    return EDLGT_OK;
  }

  // G (nv x nw) = Vp[:, 0:nv]^H X[:, 0:nw]   (X with leading dimension N)
  int gram(const S* Vp, int nv, const S* X, int nw, Dense& G) {
    G = Dense(nv, nw);
    if (nv == 0 || nw == 0) return EDLGT_OK;
    if (nw > 8) return set_error(EDLGT_ERR_INVALID_ARG, "block size above 8");
    const int nblk = (int)std::min<int64_t>(kGramBlocks, std::max<int64_t>(1, (N + 2047) / 2048));
    const int nwp = nw <= 1 ? 1 : nw <= 2 ? 2 : nw <= 4 ? 4 : 8;
    for (int j0 = 0; j0 < nv; j0 += kGramCols) {
      const int nc = std::min(kGramCols, nv - j0);
      const S* Vc = Vp + (int64_t)j0 * N;
      const unsigned fin = (unsigned)((nc * nwp + kTB / 64 - 1) / (kTB / 64));
#define GRAM_CASE(NWC)                                                                              \
  case NWC:                                                                                         \
    hipLaunchKernelGGL((gram_partial<S, NWC>), dim3(nblk), dim3(kTB), 0, st, Vc, N, nc, X, N, nw, N, \
                       part.as<double2>());                                                          \
    hipLaunchKernelGGL((gram_finish<NWC>), dim3(fin), dim3(kTB), 0, st, part.as<double2>(), nblk, nc, \
                       nw, j0, hgram.dev_as<double2>());                                                  \
    break;
      switch (nwp) { GRAM_CASE(1) GRAM_CASE(2) GRAM_CASE(4) GRAM_CASE(8) }
#undef GRAM_CASE
      EDLGT_HIP(hipGetLastError());
    }
    const double2* hbuf = hgram.as<double2>();
    EDLGT_HIP(hipStreamSynchronize(st));
    for (int j = 0; j < nv; ++j)
      for (int c = 0; c < nw; ++c) G(j, c) = cd(hbuf[(size_t)j * nw + c].x, hbuf[(size_t)j * nw + c].y);
    return EDLGT_OK;
  }

  // Reads a Gram matrix that gram_finish wrote into hgram.
  int fetch_gram(int nw, Dense& G) {
    const double2* hbuf = hgram.as<double2>();
    EDLGT_HIP(hipStreamSynchronize(st));
    G = Dense(nw, nw);
    for (int j = 0; j < nw; ++j)
      for (int c = 0; c < nw; ++c) G(j, c) = cd(hbuf[(size_t)j * nw + c].x, hbuf[(size_t)j * nw + c].y);
    return EDLGT_OK;
  }

  int upload_coef(const Dense& C, int nv, int nw) {
    std::vector<double2> hc((size_t)std::max(nv, 1) * nw);
    for (int j = 0; j < nv; ++j)
      for (int c = 0; c < nw; ++c) hc[(size_t)j * nw + c] = make_double2(C(j, c).real(), C(j, c).imag());
    EDLGT_HIP(hipMemcpyAsync(cdev.p, hc.data(), hc.size() * sizeof(double2), hipMemcpyHostToDevice, st));
    return EDLGT_OK;
  }

  // X[:, 0:nw] = beta X + Vp[:, 0:nv] C      (C: nv x nw).  With gout the
  // Gram matrix of the new X columns is computed in the same pass (nw <= 4).
  int update(const S* Vp, int nv, const Dense& C, S* X, int nw, int beta, Dense* gout = nullptr) {
    if (nw == 0) return EDLGT_OK;
    if (nv == 0 && beta && !gout) return EDLGT_OK;
    const size_t shmem = (size_t)std::max(nv, 1) * nw * sizeof(double2);
    if (shmem > 60000) {
      // split the coefficient rows: X = beta X + V[:, :h] C[:h] + V[:, h:] C[h:]
      const int half = nv / 2;
      Dense C1(half, nw), C2(nv - half, nw);
      for (int j = 0; j < nv; ++j)
        for (int c = 0; c < nw; ++c) (j < half ? C1(j, c) : C2(j - half, c)) = C(j, c);
      EDLGT_TRY(update(Vp, half, C1, X, nw, beta));
      return update(Vp + (int64_t)half * N, nv - half, C2, X, nw, 1, gout);
    }
    EDLGT_TRY(upload_coef(C, nv, nw));
    const unsigned g = grid_for(N, kTB, 4096);
    if (gout) {
      if (nw > 4) return set_error(EDLGT_ERR_INVALID_ARG, "fused Gram needs width <= 4");
      const int nwp = nw <= 1 ? 1 : nw <= 2 ? 2 : 4;
#define UPDG_CASE(NWC)                                                                              \
  case NWC:                                                                                         \
    hipLaunchKernelGGL((tall_update<S, NWC, true>), dim3(g), dim3(kTB), shmem, st, Vp, N, nv,        \
                       cdev.as<double2>(), nw, X, N, nw, N, beta, part.as<double2>());              \
    hipLaunchKernelGGL((gram_finish<NWC>), dim3((NWC * NWC + 3) / 4), dim3(kTB), 0, st,              \
                       part.as<double2>(), (int)g, NWC, nw, 0, hgram.dev_as<double2>());                 \
    break;
      switch (nwp) { UPDG_CASE(1) UPDG_CASE(2) UPDG_CASE(4) }
#undef UPDG_CASE
      EDLGT_HIP(hipGetLastError());
      return fetch_gram(nw, *gout);  // synchronises: cdev is free again
    }
#define UPD_CASE(NWC)                                                                             \
  case NWC:                                                                                       \
    hipLaunchKernelGGL((tall_update<S, NWC, false>), dim3(g), dim3(kTB), shmem, st, Vp, N, nv,     \
                       cdev.as<double2>(), nw, X, N, nw, N, beta, (double2*)nullptr);             \
    break;
    if (nw > 16) return set_error(EDLGT_ERR_INVALID_ARG, "update width above 16");
    int nwp = nw <= 1 ? 1 : nw <= 2 ? 2 : nw <= 4 ? 4 : nw <= 8 ? 8 : 16;
    switch (nwp) { UPD_CASE(1) UPD_CASE(2) UPD_CASE(4) UPD_CASE(8) UPD_CASE(16) }
#undef UPD_CASE
    EDLGT_HIP(hipGetLastError());
    EDLGT_HIP(hipStreamSynchronize(st));  // cdev is reused by the next call
    return EDLGT_OK;
  }

  // In place X = X C (C: nw x nw, nw <= 4); with gout also the new Gram matrix.
  int transform(S* X, int nw, const Dense& C, Dense* gout) {
    if (nw > 4) return set_error(EDLGT_ERR_INVALID_ARG, "transform width above 4");
    EDLGT_TRY(upload_coef(C, nw, nw));
    const unsigned g = grid_for(N, kTB, 4096);
    const int nwp = nw <= 1 ? 1 : nw <= 2 ? 2 : 4;
#define TRF_CASE(NWC)                                                                               \
  case NWC:                                                                                         \
    if (gout) {                                                                                     \
      hipLaunchKernelGGL((row_transform<S, NWC, true>), dim3(g), dim3(kTB), 0, st, X, N, nw,         \
                         cdev.as<double2>(), N, part.as<double2>());                                \
      hipLaunchKernelGGL((gram_finish<NWC>), dim3((NWC * NWC + 3) / 4), dim3(kTB), 0, st,            \
                         part.as<double2>(), (int)g, NWC, nw, 0, hgram.dev_as<double2>());               \
    } else {                                                                                        \
      hipLaunchKernelGGL((row_transform<S, NWC, false>), dim3(g), dim3(kTB), 0, st, X, N, nw,        \
                         cdev.as<double2>(), N, (double2*)nullptr);                                 \
    }                                                                                               \
    break;
    switch (nwp) { TRF_CASE(1) TRF_CASE(2) TRF_CASE(4) }
#undef TRF_CASE
    EDLGT_HIP(hipGetLastError());
    if (gout) return fetch_gram(nw, *gout);
    EDLGT_HIP(hipStreamSynchronize(st));
    return EDLGT_OK;
  }

  // In place w = c2 w + c1 b + c0 a; nrm2 = |w|^2 after the update.
  int three_term_update(const S* a, const S* bv, S* w, double c0, double c1, double c2, double& nrm2) {
    const unsigned g = grid_for(N, kTB, 4096);
    hipLaunchKernelGGL(three_term<S>, dim3(g), dim3(kTB), 0, st, a, bv, w, c0, c1, c2, N, part.as<double2>());
    hipLaunchKernelGGL((gram_finish<1>), dim3(1), dim3(kTB), 0, st, part.as<double2>(), (int)g, 1, 1, 0,
                       hgram.dev_as<double2>());
    EDLGT_HIP(hipGetLastError());
    Dense G;
    EDLGT_TRY(fetch_gram(1, G));
    nrm2 = G(0, 0).real();
    return EDLGT_OK;
  }

  int matvec(const S* X, S* Y, int nw) {
    const int vt = sizeof(S) == 16 ? EDLGT_VALUE_C128 : EDLGT_VALUE_F64;
    if (precision == EDLGT_PREC_MIXED) {
      const int64_t ncomp = (int64_t)nw * N * (sizeof(S) / 8);
      EDLGT_TRY(convert_f64_to_f32(reinterpret_cast<const double*>(X), shadow.as<float>(), ncomp, st));
      const int lt = sizeof(S) == 16 ? EDLGT_VALUE_C64 : EDLGT_VALUE_F32;
      EDLGT_TRY(csr_spmm(h, EDLGT_PREC_MIXED, lt, nw, shadow.p, N, Y, N, st));
      mv32 += nw;
    } else {
      if (nw > 1 && !pack.p && spmm_packable(h, vt)) EDLGT_TRY(pack.alloc(spmm_pack_bytes(h, vt)));
      EDLGT_TRY(csr_spmm(h, EDLGT_PREC_FP64, vt, nw, X, N, Y, N, st, pack.p, pack.bytes));
      mv64 += nw;
    }
    return EDLGT_OK;
  }

  int randomize(S* X, int64_t count) {
    hipLaunchKernelGGL(fill_random<S>, dim3(grid_for(count, 256, 4096)), dim3(256), 0, st, X, count, seed++);
    EDLGT_HIP(hipGetLastError());
    return EDLGT_OK;
  }

  // Orthonormalises X (nw columns) against V[:, :nv] and within itself.
  // On return X = Q, and Rout (nw x nw) with (X_in - V G) = Q Rout; columns
  // that were (numerically) dependent are replaced by random directions and
  // get zero rows in Rout.  Gout receives the V-coefficients.
  //
  // local > 0 (X directly follows V, nw <= 4): X = H V[:, nv-b:nv] of a
  // Lanczos step whose exact coefficients are confined to the last `local`
  // columns (block tridiagonal T).  Those are removed first (a Gram and an
  // update over local columns only), then a full Gram over V measures what
  // rounding left along the older columns, and the full update runs only
  // where that exceeds kSkipLoss.  The full pass cancels almost nothing, so
  // the DGKS test below rarely asks for another: about one sweep over V per
  // step instead of four.
  int orthonormalize(int nv, S* X, int nw, Dense& Gout, Dense& Rout, int local = 0) {
    std::vector<double> n0(nw);
    Dense Gq;  // Gram matrix of X after the Gram-Schmidt passes
    Gout = Dense(nv, nw);
    if (nv == 0) {
      EDLGT_TRY(gram(X, nw, X, nw, Gq));
      for (int c = 0; c < nw; ++c) n0[c] = std::max(Gq(c, c).real(), 0.0);
    } else {
      const S* Vp = V.as<S>();
      const bool follows = X == Vp + (int64_t)nv * N;
      const int nl = (local > 0 && local < nv && nw <= 4 && follows) ? local : 0;
      const int g0 = nv - (nl > 0 ? nl : nv);  // first column of pass 1
      const int np = nv - g0;
      // pass 1: when X directly follows V, one Gram over [V[:, g0:nv] | X]
      // gives both the coefficients and X^H X
      Dense G(np, nw), G0(nw, nw);
      if (follows) {
        Dense GG;
        EDLGT_TRY(gram(Vp + (int64_t)g0 * N, np + nw, X, nw, GG));
        for (int c = 0; c < nw; ++c) {
          for (int j = 0; j < np; ++j) G(j, c) = GG(j, c);
          for (int i = 0; i < nw; ++i) G0(i, c) = GG(np + i, c);
        }
      } else {
        EDLGT_TRY(gram(X, nw, X, nw, G0));
        EDLGT_TRY(gram(Vp, nv, X, nw, G));
      }
      for (int c = 0; c < nw; ++c) n0[c] = std::max(G0(c, c).real(), 0.0);
      Dense mG = G;
      for (auto& z : mG.a) z = -z;
      for (int j = 0; j < np; ++j)
        for (int c = 0; c < nw; ++c) Gout(g0 + j, c) = G(j, c);
      // Second pass only where the last pass lost more than half of a
      // column's squared norm (Daniel-Gragg-Kaufman-Stewart criterion): below
      // that the cancellation leaves the result orthogonal to V to working
      // precision.
      bool again = true;
      if (nw <= 4) {
        Dense G1;
        EDLGT_TRY(update(Vp + (int64_t)g0 * N, np, mG, X, nw, 1, &G1));
        std::vector<double> nref(n0);
        if (nl > 0 && local_only) {
          // no full pass: the loss of orthogonality to older columns is left
          // to grow (plain block Lanczos)
          for (int c = 0; c < nw; ++c) nref[c] = std::max(G1(c, c).real(), 0.0);
          ++ortho_skipped;
        } else if (nl > 0) {
          for (int c = 0; c < nw; ++c) nref[c] = std::max(G1(c, c).real(), 0.0);
          EDLGT_TRY(gram(Vp, nv, X, nw, G));
          // The Gram measures X's loss of orthogonality to the whole basis.
          // Below kSkipLoss (relative to each column's norm) the update is
          // skipped: X keeps those components, T keeps the exact local
          // coefficients (H V = V T + Q R stays exact), and the basis stays
          // orthogonal to kSkipLoss times one step's growth, because every
          // step measures again and corrects once the loss exceeds it.
          bool small = true;
          for (int c = 0; c < nw && small; ++c) {
            const double lim = kSkipLoss * kSkipLoss * nref[c];
            for (int j = 0; j < nv; ++j)
              if (!(std::norm(G(j, c)) <= lim)) {
                small = false;
                break;
              }
          }
          if (small) {
            ++ortho_skipped;
          } else {
            mG = G;
            for (auto& z : mG.a) z = -z;
            EDLGT_TRY(update(Vp, nv, mG, X, nw, 1, &G1));
            for (size_t i = 0; i < G.a.size(); ++i) Gout.a[i] += G.a[i];
          }
        }
        again = false;
        for (int c = 0; c < nw; ++c)
          if (!(G1(c, c).real() >= kDgks * nref[c])) again = true;
        if (!again) {
          Gq = G1;
          ++ortho_single;
        }
      } else {
        EDLGT_TRY(update(Vp, nv, mG, X, nw, 1));
      }
      if (again) {
        // pass 2, with the Gram matrix of the result formed in the same sweep
        EDLGT_TRY(gram(V.as<S>(), nv, X, nw, G));
        mG = G;
        for (auto& z : mG.a) z = -z;
        if (nw <= 4) {
          EDLGT_TRY(update(V.as<S>(), nv, mG, X, nw, 1, &Gq));
        } else {
          EDLGT_TRY(update(V.as<S>(), nv, mG, X, nw, 1));
          EDLGT_TRY(gram(X, nw, X, nw, Gq));
        }
        for (size_t i = 0; i < G.a.size(); ++i) Gout.a[i] += G.a[i];
        ++ortho_double;
      }
    }
    // Cholesky QR, twice (in place: X <- X R^-1).
    std::vector<double> floor2(nw);
    for (int c = 0; c < nw; ++c) floor2[c] = 1e-20 * std::max(n0[c], 1e-280);
    Dense R1;
    Dense Rpre(nw, nw);
    for (int c = 0; c < nw; ++c) Rpre(c, c) = 1.0;
    if (nw <= 4 && cholesky_upper(Gq, R1, floor2)) {
      if (nw == 1) {
        // one column: its squared norm from the fused Gram is exact to
        // rounding, so one scaling normalises it (a second would change it
        // by a factor 1 + O(eps))
        EDLGT_TRY(transform(X, 1, upper_inverse(R1), nullptr));
        Rout = R1;
        return EDLGT_OK;
      }
      Dense Gq2, R2;
      EDLGT_TRY(transform(X, nw, upper_inverse(R1), &Gq2));
      Rpre = R1;
      for (int c = 0; c < nw; ++c) n0[c] = 1.0;  // X now has unit-scale columns
      // The Gram of the result, formed in the same sweep, says whether the
      // first factorisation already left orthonormal columns (to kSkipLoss);
      // the second pass is only for an ill-conditioned block.
      double dev = 0.0;
      for (int i = 0; i < nw; ++i)
        for (int c = 0; c < nw; ++c) dev = std::max(dev, std::abs(Gq2(i, c) - (i == c ? 1.0 : 0.0)));
      if (dev <= kSkipLoss) {
        Rout = R1;
        return EDLGT_OK;
      }
      std::vector<double> fl2(nw, 0.25);
      if (cholesky_upper(Gq2, R2, fl2)) {
        EDLGT_TRY(transform(X, nw, upper_inverse(R2), nullptr));
        Rout = matmul(R2, R1);
        return EDLGT_OK;
      }
    }
    // Column-by-column repair (rank loss).
    Rout = Dense(nw, nw);
    for (int c = 0; c < nw; ++c) {
      S* xc = X + (int64_t)c * N;
      // against previously accepted columns of X
      for (int pass = 0; pass < 2 && c > 0; ++pass) {
        Dense g;
        EDLGT_TRY(gram(X, c, xc, 1, g));
        Dense mg = g;
        for (auto& z : mg.a) z = -z;
        EDLGT_TRY(update(X, c, mg, xc, 1, 1));
        for (int i = 0; i < c; ++i) Rout(i, c) += g(i, 0);
      }
      Dense nn;
      EDLGT_TRY(gram(xc, 1, xc, 1, nn));
      const double nrm = std::sqrt(std::max(nn(0, 0).real(), 0.0));
      if (nrm > 1e-10 * std::sqrt(std::max(n0[c], 1e-280))) {
        Dense sc(1, 1);
        sc(0, 0) = 1.0 / nrm;
        EDLGT_TRY(transform(xc, 1, sc, nullptr));
        Rout(c, c) = nrm;
      } else {
        if (nv + c + 1 > N) return set_error(EDLGT_ERR_INVALID_ARG, "basis exceeds the space");
        EDLGT_TRY(randomize(xc, N));
        for (int pass = 0; pass < 2; ++pass) {
          Dense g;
          if (nv > 0) {
            EDLGT_TRY(gram(V.as<S>(), nv, xc, 1, g));
            for (auto& z : g.a) z = -z;
            EDLGT_TRY(update(V.as<S>(), nv, g, xc, 1, 1));
          }
          if (c > 0) {
            EDLGT_TRY(gram(X, c, xc, 1, g));
            for (auto& z : g.a) z = -z;
            EDLGT_TRY(update(X, c, g, xc, 1, 1));
          }
        }
        EDLGT_TRY(gram(xc, 1, xc, 1, nn));
        Dense sc(1, 1);
        sc(0, 0) = 1.0 / std::sqrt(nn(0, 0).real());
        EDLGT_TRY(transform(xc, 1, sc, nullptr));
        Rout(c, c) = 0.0;
      }
    }
    Rout = matmul(Rout, Rpre);
    return EDLGT_OK;
  }

  // ---- resumable thick-restart state (one expansion step per matvec)
  int st_nv = 0, st_l = 0, st_k = 0;
  double st_tol = 0.0;
  int64_t st_max_mv = 0;
  Dense st_T;

  // Orthonormalises the start block V[:, 0:b] and resets the state.
  int begin(int k, double tol, int64_t max_mv) {
    Dense Gd, R;
    EDLGT_TRY(orthonormalize(0, col(0), b, Gd, R));
    st_nv = b;
    st_l = 0;
    st_k = k;
    st_tol = tol;
    st_max_mv = max_mv;
    st_steps = 0;
    st_T = Dense(M + b, M + b);
    return EDLGT_OK;
  }

  // The block the next matvec must multiply, V[:, nv-b : nv], and where its
  // product goes, V[:, nv : nv+b] (the basis has room for M + b columns).
  S* active_block() { return col(st_nv - b); }
  S* out_block() { return col(st_nv); }

  // Consumes out_block() = H * active_block(): orthogonalisation, T update, and when the
  // basis is full Rayleigh-Ritz + convergence test + thick restart.  With
  // early_every > 0 the convergence test also runs every early_every block
  // steps before the basis is full (Rayleigh-Ritz of the projected matrix on
  // the host; the residual of every Ritz pair is exact at any step:
  // H V = V T + Q Rb E^T).  Sets finished when the wanted pairs converged (or
  // the budget ran out); then theta/est hold the results and, with
  // want_vectors, Tmp[:, 0:k] the Ritz vectors.
  int early_every = 0;
  int st_steps = 0;
  int after_matvec(bool& finished, std::vector<double>& theta, std::vector<double>& est,
                   bool& converged) {
    finished = false;
    const int k = st_k;
    int& nv = st_nv;
    int& l = st_l;
    Dense& T = st_T;
    const int a0 = nv - b;
    S* Wp = out_block();
    Dense G, Rb;
    if (nv >= N) {
      // the basis spans the whole space: T = V^H H V is exact, no residual
      EDLGT_TRY(gram(V.as<S>(), nv, Wp, b, G));
      Rb = Dense(b, b);
    } else {
      // the residual block Q is orthonormalised in place: V[:, nv:nv+b].
      // Past the first block after a (re)start T is block tridiagonal, so
      // the exact coefficients sit in the last two blocks; the first block
      // after a thick restart also couples to every kept Ritz vector.
      const int local = (nv - st_l >= 2 * b) ? 2 * b : 0;
      EDLGT_TRY(orthonormalize(nv, Wp, b, G, Rb, local));
    }
    for (int i = 0; i < nv; ++i)
      for (int c = 0; c < b; ++c) {
        T(i, a0 + c) = G(i, c);
        T(a0 + c, i) = std::conj(G(i, c));
      }
    const bool full = nv + b > M || nv + b > N;
    if (!full) {
      for (int c = 0; c < b; ++c)
        for (int i = 0; i < b; ++i) {
          T(nv + i, a0 + c) = Rb(i, c);
          T(a0 + c, nv + i) = std::conj(Rb(i, c));
        }
    }
    ++st_steps;
    const bool test = full || (early_every > 0 && nv >= k && st_steps % early_every == 0);
    if (!test) {
      nv += b;
      return EDLGT_OK;
    }
    // Rayleigh-Ritz
    Dense Tn(nv, nv);
    for (int i = 0; i < nv; ++i)
      for (int j = 0; j < nv; ++j) Tn(i, j) = T(i, j);
    std::vector<double> w;
    Dense Sv;
    hermitian_eig(Tn, w, Sv);
    const bool exhausted = nv >= N;
    theta.assign(w.begin(), w.begin() + k);
    est.assign(k, 0.0);
    bool ok = true;
    for (int i = 0; i < k; ++i) {
      double s2 = 0.0;
      if (!exhausted) {
        for (int c = 0; c < b; ++c) {
          cd z = 0.0;
          for (int q = 0; q < b; ++q) z += Rb(c, q) * Sv(a0 + q, i);
          s2 += std::norm(z);
        }
      }
      est[i] = std::sqrt(s2);
      if (est[i] > st_tol * std::max(1.0, std::fabs(w[i]))) ok = false;
    }
    const int64_t mv = mv64 + mv32;
    if (ok || exhausted || mv >= st_max_mv) {
      converged = ok || exhausted;
      finished = true;
      if (want_vectors) {
        Dense Sk(nv, k);
        for (int j = 0; j < nv; ++j)
          for (int i = 0; i < k; ++i) Sk(j, i) = Sv(j, i);
        EDLGT_TRY(update_blocked(col(0), nv, Sk, Tmp.as<S>(), k));
      }
      return EDLGT_OK;
    }
    if (!full) {
      nv += b;
      return EDLGT_OK;
    }
    // thick restart: keep the lowest l Ritz vectors and the residual block
    l = std::max(k, std::min(M - 2 * b, k + (M - k) / 2));
    l = std::min(l, nv - b);
    Dense Sl(nv, l);
    for (int j = 0; j < nv; ++j)
      for (int i = 0; i < l; ++i) Sl(j, i) = Sv(j, i);
    // Ritz vectors into Tmp[:, 0:l], the residual block after them, then Tmp
    // becomes the basis (both buffers hold M + b columns)
    EDLGT_TRY(update_blocked(col(0), nv, Sl, Tmp.as<S>(), l));
    EDLGT_HIP(hipMemcpyAsync(Tmp.as<S>() + (int64_t)l * N, col(nv), (size_t)b * N * sizeof(S),
                             hipMemcpyDeviceToDevice, st));
    std::swap(V.p, Tmp.p);
    std::swap(V.bytes, Tmp.bytes);
    for (auto& z : T.a) z = 0.0;
    for (int i = 0; i < l; ++i) T(i, i) = w[i];
    for (int c = 0; c < b; ++c)
      for (int i = 0; i < l; ++i) {
        cd z = 0.0;
        for (int q = 0; q < b; ++q) z += Rb(c, q) * Sv(a0 + q, i);
        T(l + c, i) = z;
        T(i, l + c) = std::conj(z);
      }
    nv = l + b;
    return EDLGT_OK;
  }

  // Runs thick-restart block Lanczos from the b start columns already in
  // V[:, 0:b] (not necessarily orthonormal).  Fills theta (k values), and the
  // Ritz vectors Y (k columns) into Tmp[:, 0:k]; est gets the estimates.
  int run(int k, double tol, int64_t max_mv, std::vector<double>& theta, std::vector<double>& est,
          bool& converged) {
    EDLGT_TRY(begin(k, tol, max_mv));
    converged = false;
    while (true) {
      EDLGT_TRY(matvec(active_block(), out_block(), b));
      bool finished = false;
      EDLGT_TRY(after_matvec(finished, theta, est, converged));
      if (finished) return EDLGT_OK;
    }
  }

  // X[:, 0:nc] = V[:, 0:nv] C for any nc (chunks of up to 16 columns: V is
  // read once per chunk; a 32-column kernel holds twice the accumulators and
  // measured slower than two 16-column passes).
  int update_blocked(const S* Vp, int nv, const Dense& C, S* X, int nc) {
    for (int c0 = 0; c0 < nc; c0 += 16) {
      const int w = std::min(16, nc - c0);
      Dense Cc(nv, w);
      for (int j = 0; j < nv; ++j)
        for (int c = 0; c < w; ++c) Cc(j, c) = C(j, c0 + c);
      EDLGT_TRY(update(Vp, nv, Cc, X + (int64_t)c0 * N, w, 0));
    }
    return EDLGT_OK;
  }

  // a-posteriori error of the Krylov exponential with the first k basis
  // vectors: beta_k |(e^{-i T_k dt} e_1)_k| (T_k the leading k x k block)
  static double krylov_estimate(const std::vector<double>& alpha, const std::vector<double>& beta, int k,
                                double dt) {
    Dense T(k, k);
    for (int j = 0; j < k; ++j) {
      T(j, j) = alpha[j];
      if (j + 1 < k) {
        T(j, j + 1) = beta[j];
        T(j + 1, j) = beta[j];
      }
    }
    std::vector<double> lam;
    Dense Sv;
    hermitian_eig(T, lam, Sv);
    cd z = 0.0;
    for (int l = 0; l < k; ++l) z += Sv(k - 1, l) * std::exp(cd(0.0, -lam[l] * dt)) * std::conj(Sv(0, l));
    return beta[k - 1] * std::abs(z);
  }

  // One Krylov step psi <- exp(-i H dt) psi (Lanczos basis of at most m
  // vectors, three-term recurrence plus optional full CGS2
  // reorthogonalisation).  V[:, 0..m] holds the basis; the projected
  // exponential is formed on the host from the eigendecomposition of the real
  // tridiagonal T.  The recurrence stops early once the standard
  // a-posteriori estimate beta_k |(e^{-iT_k dt} e_1)_k| (relative to |psi|)
  // is below kKrylovTol, tested from kKrylovMin vectors on; err receives that
  // estimate for the basis used (0 for an invariant subspace).
  int krylov_step(S* psi, double dt, int m, bool reorth, double& err, int& m_used) {
    Dense nn;
    EDLGT_TRY(gram(psi, 1, psi, 1, nn));
    const double beta0 = std::sqrt(std::max(nn(0, 0).real(), 0.0));
    err = 0.0;
    if (beta0 == 0.0) {
      m_used = 0;
      return EDLGT_OK;
    }
    Dense sc(1, 1);
    sc(0, 0) = 1.0 / beta0;
    EDLGT_TRY(update(psi, 1, sc, col(0), 1, 0));
    std::vector<double> alpha, beta;
    // Without reorthogonalisation the basis is stored unnormalised, u_j =
    // nrm[j] v_j: the product lands in column j+1 and one fused pass forms
    // u_{j+1} = H v_j - alpha_j v_j - beta_{j-1} v_{j-1} and its norm.
    std::vector<double> nrm(1, 1.0);
    double hnorm = 0.0;
    int me = m;
    bool exact = false;
    for (int j = 0; j < m; ++j) {
      double a = 0.0, b = 0.0;
      if (reorth) {
        EDLGT_TRY(matvec(col(j), W.as<S>(), 1));
        for (int pass = 0; pass < 2; ++pass) {
          Dense G;
          EDLGT_TRY(gram(V.as<S>(), j + 1, W.as<S>(), 1, G));
          a += G(j, 0).real();
          for (auto& z : G.a) z = -z;
          EDLGT_TRY(update(V.as<S>(), j + 1, G, W.as<S>(), 1, 1));
        }
        Dense bn;
        EDLGT_TRY(gram(W.as<S>(), 1, W.as<S>(), 1, bn));
        b = std::sqrt(std::max(bn(0, 0).real(), 0.0));
      } else {
        S* u = col(j);
        S* w = col(j + 1);
        EDLGT_TRY(matvec(u, w, 1));
        Dense G;
        EDLGT_TRY(gram(u, 1, w, 1, G));
        const double nj = nrm[j];
        a = G(0, 0).real() / (nj * nj);
        const double cp = j > 0 ? -beta[j - 1] / nrm[j - 1] : 0.0;
        double nn = 0.0;
        EDLGT_TRY(three_term_update(j > 0 ? col(j - 1) : u, u, w, cp, -a / nj, 1.0 / nj, nn));
        b = std::sqrt(std::max(nn, 0.0));
      }
      alpha.push_back(a);
      beta.push_back(b);
      hnorm = std::max(hnorm, std::fabs(a) + b + (j > 0 ? beta[j - 1] : 0.0));
      if (b <= 1e-14 * std::max(hnorm, 1e-300)) {  // invariant subspace: exact
        me = j + 1;
        exact = true;
        break;
      }
      if (j + 1 >= kKrylovMin && j + 1 < m && krylov_estimate(alpha, beta, j + 1, dt) <= kKrylovTol) {
        me = j + 1;
        break;
      }
      if (reorth) {
        Dense s1(1, 1);
        s1(0, 0) = 1.0 / b;
        EDLGT_TRY(update(W.as<S>(), 1, s1, col(j + 1), 1, 0));
        nrm.push_back(1.0);
      } else {
        nrm.push_back(b);
      }
    }
    Dense T(me, me);
    for (int j = 0; j < me; ++j) {
      T(j, j) = alpha[j];
      if (j + 1 < me) {
        T(j, j + 1) = beta[j];
        T(j + 1, j) = beta[j];
      }
    }
    std::vector<double> lam;
    Dense Sv;
    hermitian_eig(T, lam, Sv);
    Dense c(me, 1);
    for (int k = 0; k < me; ++k) {
      cd z = 0.0;
      for (int l = 0; l < me; ++l) z += Sv(k, l) * std::exp(cd(0.0, -lam[l] * dt)) * std::conj(Sv(0, l));
      c(k, 0) = beta0 * z;
    }
    err = exact ? 0.0 : beta[me - 1] * std::abs(c(me - 1, 0)) / beta0;
    for (int k = 0; k < me; ++k) c(k, 0) /= nrm[k];
    EDLGT_TRY(update(V.as<S>(), me, c, psi, 1, 0));
    m_used = me;
    return EDLGT_OK;
  }

  // True residual norms ||H y_i - theta_i y_i|| in FP64 for Y = Tmp[:, 0:k].
  // Inexact Newton refinement of the lowest eigenpair in place (x: unit
  // FP64 vector, e.g. the Ritz vector of the FP32 phase), for the mixed
  // precision mode: each outer step forms the FP64 residual r = H x - theta x
  // (theta the Rayleigh quotient) and solves the projected correction equation
  // (P (H32 - theta) P) delta = -r, delta orthogonal to x, by conjugate
  // gradients with the FP32 matrix to a relative residual eta; then
  // x <- (x + delta) / |x + delta|.  On the complement of x, H - theta is
  // positive definite with smallest eigenvalue lambda_1 - lambda_0, so CG
  // converges at the rate Lanczos does, but needs no basis and no
  // reorthogonalisation.  ok = false (x left as a start vector for the FP64
  // Lanczos phase) when the residual stops falling or the budget is spent.
  int refine_newton(S* x, double tol, double eta, int64_t max_mv, double& theta, bool& ok) {
    ok = false;
    if (M + b < 5 || !shadow.p || h.mf) return EDLGT_OK;
    using L = typename std::conditional<sizeof(S) == 16, float2, float>::type;
    // FP64 columns 0-1 hold H x and the FP32 products; columns 2-4 each hold
    // two FP32 vectors
    S *hx = col(0), *y = col(1);
    L* xlo = reinterpret_cast<L*>(col(2));
    L *g = xlo + N, *r = reinterpret_cast<L*>(col(3)), *pp = r + N, *delta = reinterpret_cast<L*>(col(4)),
      *q = delta + N;
    double2* pg = part.as<double2>();
    double2* ppq = pg + kCgBlocks;
    double2* prr[2] = {pg + 2 * kCgBlocks, pg + 3 * kCgBlocks};
    const int lt = sizeof(S) == 16 ? EDLGT_VALUE_C64 : EDLGT_VALUE_F32;
    auto read_sum = [&](const double2* pr, double& out) -> int {
      EDLGT_HIP(hipGetLastError());
      hipLaunchKernelGGL((gram_finish<1>), dim3(1), dim3(kTB), 0, st, pr, kCgBlocks, 1, 1, 0,
                         hgram.dev_as<double2>());
      EDLGT_HIP(hipGetLastError());
      Dense G;
      EDLGT_TRY(fetch_gram(1, G));
      out = G(0, 0).real();
      return EDLGT_OK;
    };
    const int saved = precision;
    double prev = INFINITY;
    int64_t used = 0;
    for (int outer = 0; outer < 6; ++outer) {
      // FP64 residual of the current vector
      precision = EDLGT_PREC_FP64;
      EDLGT_TRY(matvec(x, hx, 1));
      ++used;
      Dense xh;
      EDLGT_TRY(gram(x, 1, hx, 1, xh));
      theta = xh(0, 0).real();
      hipLaunchKernelGGL((resid_rr<S>), dim3(kCgBlocks), dim3(kTB), 0, st, hx, x, theta, N, prr[0]);
      double rn = 0.0;
      EDLGT_TRY(read_sum(prr[0], rn));
      const double res = std::sqrt(std::max(rn, 0.0));
      if (res <= tol * std::max(1.0, std::fabs(theta))) {
        ok = true;
        break;
      }
      if (!(res < 0.5 * prev) || used >= max_mv) break;
      prev = res;
      // y = H32 x_lo; g = y - theta x; r = -(H x - theta x); delta = 0
      const int64_t ncomp = N * (int64_t)(sizeof(S) / 8);
      EDLGT_TRY(convert_f64_to_f32(reinterpret_cast<const double*>(x), reinterpret_cast<float*>(xlo), ncomp, st));
      EDLGT_TRY(csr_spmm(h, EDLGT_PREC_MIXED, lt, 1, xlo, N, y, N, st));
      ++used;
      ++mv32;
      hipLaunchKernelGGL((cg_init<S, L>), dim3(kCgBlocks), dim3(kTB), 0, st, y, hx, x, theta, N, g, r, delta,
                         prr[0]);
      double rr0 = 0.0;
      EDLGT_TRY(read_sum(prr[0], rr0));
      const double target = eta * eta * rr0;
      bool inner_ok = false;
      double hist[4] = {rr0, rr0, rr0, rr0};  // |r|^2 at the last four checks
      for (int it = 0; it < kCgMaxIter && used < max_mv; ++it) {
        hipLaunchKernelGGL((cg_dir<L>), dim3(kCgBlocks), dim3(kTB), 0, st, r, pp, g, prr[it & 1],
                           prr[(it + 1) & 1], it == 0 ? 1 : 0, N, pg);
        EDLGT_HIP(hipGetLastError());
        EDLGT_TRY(csr_spmm(h, EDLGT_PREC_MIXED, lt, 1, pp, N, y, N, st));
        ++used;
        ++mv32;
        hipLaunchKernelGGL((cg_apply<S, L>), dim3(kCgBlocks), dim3(kTB), 0, st, pp, y, xlo, q, pg, theta, N, ppq);
        hipLaunchKernelGGL((cg_step<L>), dim3(kCgBlocks), dim3(kTB), 0, st, delta, r, pp, q, prr[it & 1], ppq, N,
                           prr[(it + 1) & 1]);
        EDLGT_HIP(hipGetLastError());
        if (it % 4 == 3) {
          double rr = 0.0;
          EDLGT_TRY(read_sum(prr[(it + 1) & 1], rr));
          if (!(rr <= 1e8 * rr0)) break;  // diverging (or NaN): not positive definite
          if (rr <= target) {
            inner_ok = true;
            break;
          }
          // stalled at the FP32 floor (|r| fell less than 2x over 16
          // iterations): keep the correction if it reduced |r| 100-fold
          if (it >= 15 && rr > 0.25 * hist[(it / 4) & 3]) {
            inner_ok = rr <= 1e-4 * rr0;
            break;
          }
          hist[(it / 4) & 3] = rr;
        }
      }
      if (!inner_ok) break;
      // x <- (x + delta) / |x + delta|
      hipLaunchKernelGGL((cg_accept<S, L>), dim3(kCgBlocks), dim3(kTB), 0, st, x, delta, N, prr[0]);
      double xn = 0.0;
      EDLGT_TRY(read_sum(prr[0], xn));
      Dense sc(1, 1);
      sc(0, 0) = 1.0 / std::sqrt(std::max(xn, 1e-300));
      EDLGT_TRY(transform(x, 1, sc, nullptr));
    }
    precision = saved;
    return EDLGT_OK;
  }

  int residuals(int k, const std::vector<double>& theta, std::vector<double>& out) {
    out.assign(k, 0.0);
    const int saved = precision;
    precision = EDLGT_PREC_FP64;
    for (int i = 0; i < k; ++i) {
      S* y = Tmp.as<S>() + (int64_t)i * N;
      EDLGT_TRY(matvec(y, W.as<S>(), 1));
      Dense c(1, 1);
      c(0, 0) = -theta[i];
      EDLGT_TRY(update(y, 1, c, W.as<S>(), 1, 1));
      Dense nn;
      EDLGT_TRY(gram(W.as<S>(), 1, W.as<S>(), 1, nn));
      out[i] = std::sqrt(std::max(nn(0, 0).real(), 0.0));
    }
    precision = saved;
    return EDLGT_OK;
  }
};

// Basis capacity M (a multiple of the block size b) for k wanted pairs.  A
// basis that can hold the whole space (N <= 4096) switches to single vectors
// run to exhaustion: always for N <= 512 when M >= N, and for larger N when
// the half-sector cap would leave fewer than k + 2b vectors (a thick restart
// keeps k + b of them and needs room for one more block, else the highest
// wanted pairs are dropped at every restart and the solve cannot converge).
// Above 4096 states such a k is an error: the projected matrix would have N
// rows and is factored on the host.
int choose_basis(int64_t N, int k, int max_basis, int& b, int& M) {
  M = std::max(max_basis, 2 * k + 2 * b);
  M = (M + b - 1) / b * b;
  if (N <= 512 && M >= N) {
    b = 1;
    M = (int)N;
    return EDLGT_OK;
  }
  if (M > N / 2) {
    M = (int)std::max<int64_t>(2 * b, (N / 2) / b * b);
    if (M < k + 2 * b) {
      if (N > 4096)
        return set_error(EDLGT_ERR_INVALID_ARG,
                         "n_eigs too large for the sector: a restarted basis of at most N/2 vectors "
                         "cannot hold n_eigs + 2 blocks");
      b = 1;
      M = (int)N;
    }
  }
  return EDLGT_OK;
}

template <class S>
int lanczos_typed(Csr& h, int k, double tol, int max_basis, int64_t max_iter, int precision,
                  const void* v0, double* evals, void* evecs, double* resid, int64_t* matvecs,
                  hipStream_t st) {
  const int64_t N = h.n_rows;
  if (k < 1 || k > N) return set_error(EDLGT_ERR_INVALID_ARG, "n_eigs must be in [1, n]");
  Solver<S> sv(h, st, precision);
  // Block size: up to 4 so that small degeneracies are resolved; a basis
  // that can hold the whole space uses single vectors and runs to exhaustion.
  int b = std::min(k, 4);
  // max_basis <= 0: ARPACK's default basis size ncv = max(2k + 1, 20), the
  // size the reference's eigsh uses (measured the faster choice for one
  // eigenvalue: the reorthogonalisation cost grows with the basis), and at
  // least 16 blocks when the block holds several vectors (k = 4 lowest of
  // the Z3 4x4 sector: 5 blocks of 4 needed 2008 matvecs; 16 blocks were the
  // fastest of 5 to 24 on the served couplings)
  if (max_basis <= 0) max_basis = std::max(std::max(2 * k + 1, 20), b > 1 ? 16 * b : 0);
  int M = 0;
  EDLGT_TRY(choose_basis(N, k, max_basis, b, M));
  sv.b = b;
  sv.M = M;
  // convergence tested after every block step (a host Rayleigh-Ritz of the
  // projected matrix, at most M x M), not only when the basis is full: the
  // solve stops within one step of convergence instead of up to a basis later
  if (M < N && M <= 128) sv.early_every = 1;
  EDLGT_TRY(sv.setup());
  // start block: caller's v0 in column 0, random columns after it
  if (v0) {
    // host or device memory (unified addressing)
    EDLGT_HIP(hipMemcpyAsync(sv.col(0), v0, (size_t)N * sizeof(S), hipMemcpyDefault, st));
  } else {
    EDLGT_TRY(sv.randomize(sv.col(0), N));
  }
  if (b > 1) EDLGT_TRY(sv.randomize(sv.col(1), (int64_t)(b - 1) * N));
  std::vector<double> theta, est;
  bool conv = false;
  bool refined = false;
  if (precision == EDLGT_PREC_MIXED && k == 1) {
    // one eigenpair: FP32-product Lanczos to a loose tolerance, then inexact
    // Newton refinement (FP64 residuals, FP32 CG corrections); the FP64
    // Lanczos phase below only if the refinement stalls
    const double tol32 = std::max(tol, env_double("EDLGT_MIXED_TOL32", kMixedTol32));
    sv.local_only = env_double("EDLGT_MIXED_LOCAL", 1.0) != 0.0;
    EDLGT_TRY(sv.run(k, tol32, max_iter, theta, est, conv));
    sv.local_only = false;
    double th = 0.0;
    EDLGT_TRY(sv.refine_newton(sv.Tmp.template as<S>(), tol, env_double("EDLGT_MIXED_ETA", kMixedEta),
                               max_iter - sv.mv64 - sv.mv32, th, refined));
    if (refined) {
      theta.assign(1, th);
      conv = true;
    }
  }
  if (precision == EDLGT_PREC_MIXED && !refined) {
    // FP32 phase to a loose tolerance, then FP64 from its Ritz block
    const double tol32 = std::max(tol, 2e-6);
    if (k != 1) EDLGT_TRY(sv.run(k, tol32, max_iter, theta, est, conv));
    const int nk = std::min(k, b);
    EDLGT_HIP(hipMemcpyAsync(sv.col(0), sv.Tmp.p, (size_t)nk * N * sizeof(S), hipMemcpyDeviceToDevice, st));
    if (nk < b) EDLGT_TRY(sv.randomize(sv.col(nk), (int64_t)(b - nk) * N));
    if (k > b) {
      // more wanted vectors than the block: fold the rest into the block
      Dense c(1, 1);
      c(0, 0) = 1.0;
      for (int i = b; i < k; ++i)
        EDLGT_TRY(sv.update(sv.Tmp.template as<S>() + (int64_t)i * N, 1, c, sv.col(i % b), 1, 1));
    }
    sv.precision = EDLGT_PREC_FP64;
  }
  if (!refined) EDLGT_TRY(sv.run(k, tol, max_iter, theta, est, conv));
  for (int i = 0; i < k; ++i) evals[i] = theta[i];
  if (evecs)
    EDLGT_HIP(hipMemcpyAsync(evecs, sv.Tmp.p, (size_t)k * N * sizeof(S), hipMemcpyDeviceToDevice, st));
  if (resid) {
    std::vector<double> r;
    EDLGT_TRY(sv.residuals(k, theta, r));
    for (int i = 0; i < k; ++i) resid[i] = r[i];
  }
  if (matvecs) {
    matvecs[0] = sv.mv64;
    matvecs[1] = sv.mv32;
  }
  EDLGT_HIP(hipStreamSynchronize(st));
  if (!conv) return set_error(EDLGT_ERR_NOT_CONVERGED, "Lanczos did not converge within max_iter");
  return EDLGT_OK;
}

}  // namespace

// Krylov stepping state that outlives one call: the basis of m+1 vectors
// (m+1 full-length vectors: 187 GB for 377M complex128 states) and the
// marginal buffer stay allocated between runs, so a trajectory computed in
// chunks pays the allocation once.
struct KrylovSession {
  Csr& h;
  const Basis* basis;
  int m, reorth, d_max, bins;
  hipStream_t st;
  std::unique_ptr<Solver<double2>> sv;
  DevBuf marg;
  KrylovSession(Csr& hh, const Basis* b, hipStream_t s) : h(hh), basis(b), st(s) {}
};

int krylov_create(Csr& h, const Basis* basis, int precision, int m, int reorth, int d_max, hipStream_t st,
                  KrylovSession** out) {
  const int64_t N = h.n_rows;
  m = (int)std::max<int64_t>(1, std::min<int64_t>(m, N));
  auto ks = std::make_unique<KrylovSession>(h, basis, st);
  ks->m = m;
  ks->reorth = reorth;
  ks->d_max = d_max;
  ks->bins = basis && d_max > 0 ? basis->n_sites * d_max : 0;
  ks->sv = std::make_unique<Solver<double2>>(h, st, precision);
  ks->sv->b = 1;
  ks->sv->M = m;
  ks->sv->need_tmp = false;
  ks->sv->need_w = reorth != 0;
  EDLGT_TRY(ks->sv->setup());
  if (ks->bins > 0) EDLGT_TRY(ks->marg.alloc((size_t)ks->bins * sizeof(double)));
  *out = ks.release();
  return EDLGT_OK;
}

void krylov_destroy(KrylovSession* ks) { delete ks; }

int krylov_session_run(KrylovSession& ks, void* psi, double dt, int64_t n_steps, double* out_marg,
                       double* out_err, void* history) {
  using S = double2;
  const int64_t N = ks.h.n_rows;
  if (out_marg && ks.bins == 0) return set_error(EDLGT_ERR_INVALID_ARG, "marginals need the sector basis");
  auto record = [&](int64_t t) -> int {
    if (out_marg) {
      EDLGT_TRY(basis_marginals(*ks.basis, psi, true, ks.d_max, ks.marg.as<double>(), ks.st));
      EDLGT_HIP(hipMemcpy(out_marg + t * ks.bins, ks.marg.p, (size_t)ks.bins * sizeof(double),
                          hipMemcpyDeviceToHost));
    }
    if (history)
      EDLGT_HIP(hipMemcpyAsync(static_cast<S*>(history) + t * N, psi, (size_t)N * sizeof(S),
                               hipMemcpyDeviceToDevice, ks.st));
    return EDLGT_OK;
  };
  EDLGT_TRY(record(0));
  for (int64_t t = 0; t < n_steps; ++t) {
    double err = 0.0;
    int mu = 0;
    EDLGT_TRY(ks.sv->krylov_step(static_cast<S*>(psi), dt, ks.m, ks.reorth != 0, err, mu));
    if (out_err) out_err[t] = err;
    EDLGT_TRY(record(t + 1));
  }
  EDLGT_HIP(hipStreamSynchronize(ks.st));
  return EDLGT_OK;
}

namespace {

int krylov_typed(Csr& h, const Basis* basis, int precision, void* psi, double dt, int64_t n_steps,
                 int m, int reorth, int d_max, double* out_marg, double* out_err, void* history,
                 hipStream_t st) {
  if (out_marg && !basis) return set_error(EDLGT_ERR_INVALID_ARG, "marginals need the sector basis");
  KrylovSession* ks = nullptr;
  EDLGT_TRY(krylov_create(h, basis, precision, m, reorth, out_marg ? d_max : 0, st, &ks));
  std::unique_ptr<KrylovSession> hold(ks);
  return krylov_session_run(*ks, psi, dt, n_steps, out_marg, out_err, history);
}

// ---------------------------------------------------------------- serving
// Continuous-batching eigensolver: up to max_slots independent requests, each
// a block thick-restart Lanczos run for the lowest k eigenvalues of
// H(c) = sum_t c[t] H_t on one shared multi-term CSR pattern.  One step
// advances every active slot by one block expansion: a single fused SpMM
// applies each slot's own coefficients to its active block (the matrix is
// streamed once for all slots), then each slot orthogonalises, updates its
// projected matrix, tests convergence and, when its basis is full, restarts.
// Finished slots are freed at once and take a new request before the next
// step.  Slots are allocated on first use (slot_bytes() each), so a caller can
// admit by device-memory footprint.
//
// Start blocks: cold (warm_cap == 0) from the request's seed only, so a
// result is bitwise the same alone or in any mix (every column of the fused
// SpMM is computed in the same order for any batch).  Warm (warm_cap > 0):
// the server keeps the Ritz vectors of up to warm_cap finished requests on the
// device and starts a new request from those of the nearest coefficient
// vector, or, when a second cached coupling lies on the other side of the
// request, from the two interpolated after a Procrustes rotation; the result
// then agrees with the cold one to the requested tolerance.
template <class S>
struct Server {
  Csr& h;
  hipStream_t st;
  int vec_type, b, M, k;
  double tol;
  int64_t max_mv;
  int max_slots = 0;
  struct Slot {
    std::unique_ptr<Solver<S>> sv;
    bool active = false;
    int64_t req = -1;
    double coef[16] = {0};
  };
  std::vector<Slot> slots;
  // per step: device pointer arrays and coefficients of every active column
  // (xs, ys, cf), staged in pinned host memory (hxp, hyp, hcp) so that a
  // captured graph's copy nodes read the current step's values
  DevBuf xs, ys, cf, pack;
  HostPinned hxp, hyp, hcp;
  int max_cols = 0;
  // the fused-SpMM chain of a step (uploads, coefficient tables, pack + SpMM
  // per group) captured once per column count and replayed (the chain's
  // launches are fixed by the column count, the matrix and the workspace)
  struct StepGraph {
    int ncol;
    void* pack;
    hipGraphExec_t exec;
  };
  std::vector<StepGraph> graphs;
  bool use_graphs = false;
  int64_t graph_launches = 0, graph_captures = 0;
  // warm-start cache
  int warm_cap = 0, warm_n = 0;
  DevBuf warm;
  std::vector<std::vector<double>> warm_coef;

  Server(Csr& hh, hipStream_t s, int vt) : h(hh), st(s), vec_type(vt) {}
  ~Server() { drop_graphs(); }
  void drop_graphs() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }

  int setup(int max_slots_, int n_eigs, double tol_, int max_basis, int64_t max_mv_, int warm_cap_) {
    const int64_t N = h.n_rows;
    k = n_eigs;
    tol = tol_;
    max_mv = max_mv_;
    max_slots = max_slots_;
    warm_cap = warm_cap_;
    b = std::min(k, 4);
    EDLGT_TRY(choose_basis(N, k, max_basis, b, M));
    slots.resize(max_slots);
    max_cols = max_slots * b;
    EDLGT_TRY(xs.alloc(max_cols * sizeof(void*)));
    EDLGT_TRY(ys.alloc(max_cols * sizeof(void*)));
    EDLGT_TRY(cf.alloc((size_t)(max_cols + kCoefMaxRhs) * 16 * sizeof(double)));
    EDLGT_TRY(hxp.alloc(max_cols * sizeof(void*)));
    EDLGT_TRY(hyp.alloc(max_cols * sizeof(void*)));
    EDLGT_TRY(hcp.alloc((size_t)(max_cols + kCoefMaxRhs) * 16 * sizeof(double)));
    std::memset(hcp.p, 0, hcp.bytes);
    // graph capture needs a stream other than the legacy default one;
    // EDLGT_SERVER_GRAPHS=0 launches the chain eagerly (same results)
    const char* env = std::getenv("EDLGT_SERVER_GRAPHS");
    use_graphs = st != nullptr && !(env && env[0] == '0');
    return EDLGT_OK;
  }

  // device bytes of one slot's solver (Solver::setup with need_w = false):
  // basis and restart buffer, Gram partials and the coefficient buffer; its
  // Gram result buffer (same size as the coefficients) is pinned host memory
  size_t slot_bytes() const {
    const size_t N = (size_t)h.n_rows;
    return ((size_t)(M + b) + std::max(M + b, 8)) * N * sizeof(S) +
           (size_t)kGramBlocks * kGramCols * 8 * sizeof(double2) +
           (size_t)(M + b) * std::max(M, 8) * sizeof(double2);
  }
  size_t slot_host_bytes() const { return (size_t)(M + b) * std::max(M, 8) * sizeof(double2); }
  // warm cache plus the packed right-hand-side workspace of the fused SpMM
  size_t shared_bytes() const {
    return (size_t)warm_cap * k * (size_t)h.n_rows * sizeof(S) + coef_pack_bytes(h, vec_type, kCoefMaxRhs);
  }

  int active() const {
    int a = 0;
    for (auto& sl : slots) a += sl.active ? 1 : 0;
    return a;
  }
  int allocated() const {
    int a = 0;
    for (auto& sl : slots) a += sl.sv ? 1 : 0;
    return a;
  }

  // Frees the memory of every idle slot.
  void trim() {
    for (auto& sl : slots)
      if (sl.sv && !sl.active) sl.sv.reset();
  }

  int nearest_warm(const double* c) const {
    int best = -1;
    double bd = 0.0;
    for (int e = 0; e < warm_n; ++e) {
      double d = 0.0;
      for (int t = 0; t < h.n_terms; ++t) d += (warm_coef[e][t] - c[t]) * (warm_coef[e][t] - c[t]);
      if (best < 0 || d < bd) {
        best = e;
        bd = d;
      }
    }
    return best;
  }

  // The cached coupling nearest to c among those with c between them and
  // entry e1 (projection weight w of c on the segment e1 -> e in (0, 1)),
  // or -1.
  int bracket_warm(const double* c, int e1, double& w) const {
    int best = -1;
    double bd = 0.0;
    for (int e = 0; e < warm_n; ++e) {
      if (e == e1) continue;
      double num = 0.0, den = 0.0, d = 0.0;
      for (int t = 0; t < h.n_terms; ++t) {
        const double s1 = warm_coef[e][t] - warm_coef[e1][t];
        num += (c[t] - warm_coef[e1][t]) * s1;
        den += s1 * s1;
        d += (warm_coef[e][t] - c[t]) * (warm_coef[e][t] - c[t]);
      }
      if (!(den > 0.0)) continue;
      const double we = num / den;
      if (!(we > 0.0 && we < 1.0)) continue;
      if (best < 0 || d < bd) {
        best = e;
        bd = d;
        w = we;
      }
    }
    return best;
  }

  // Start block between two cached couplings: the Ritz block Y2 is rotated
  // onto Y1 (orthogonal Procrustes: Q = polar(Y2^H Y1), which fixes the
  // signs, phases and the mixing inside degenerate groups) and the two are
  // interpolated, X = (1 - w) Y1 + w Y2 Q, an approximation second order in
  // the coupling step where either block alone is first order.  Leaves
  // done = false (X untouched) when the two subspaces are too far apart.
  int interpolate_warm(Solver<S>& sv, int e1, int e2, double w, bool& done) {
    done = false;
    const int64_t N = h.n_rows;
    const S* Y1 = warm.as<S>() + (size_t)e1 * k * N;
    const S* Y2 = warm.as<S>() + (size_t)e2 * k * N;
    Dense G;
    EDLGT_TRY(sv.gram(Y1, k, Y2, k, G));  // Y1^H Y2
    Dense GG(k, k);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        cd z = 0.0;
        for (int q = 0; q < k; ++q) z += G(i, q) * std::conj(G(j, q));
        GG(i, j) = z;
      }
    std::vector<double> lam;
    Dense W;
    hermitian_eig(GG, lam, W);
    if (!(lam[0] >= 0.25)) return EDLGT_OK;  // a principal angle above 60 degrees
    Dense Sm(k, k);  // (G G^H)^(-1/2)
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        cd z = 0.0;
        for (int q = 0; q < k; ++q) z += W(i, q) * (1.0 / std::sqrt(lam[q])) * std::conj(W(j, q));
        Sm(i, j) = z;
      }
    Dense Gh(k, k);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) Gh(i, j) = std::conj(G(j, i));
    const Dense Q = matmul(Gh, Sm);
    Dense C1(k, b), C2(k, b);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < b; ++j) {
        C1(i, j) = i == j ? cd(1.0 - w, 0.0) : cd(0.0, 0.0);
        C2(i, j) = w * Q(i, j);
      }
    EDLGT_TRY(sv.update(Y1, k, C1, sv.col(0), b, 0));
    EDLGT_TRY(sv.update(Y2, k, C2, sv.col(0), b, 1));
    done = true;
    return EDLGT_OK;
  }

  // Prefers an allocated free slot; allocates a new one (up to max_slots)
  // otherwise.  *slot_out = -1 when every slot is busy.
  int submit(int64_t req, const double* coef, uint64_t seed, int* slot_out) {
    *slot_out = -1;
    int pick = -1;
    for (size_t i = 0; i < slots.size() && pick < 0; ++i)
      if (slots[i].sv && !slots[i].active) pick = (int)i;
    for (size_t i = 0; i < slots.size() && pick < 0; ++i)
      if (!slots[i].sv) {
        Slot& sl = slots[i];
        sl.sv.reset(new Solver<S>(h, st, EDLGT_PREC_FP64));
        sl.sv->b = b;
        sl.sv->M = M;
        sl.sv->want_vectors = warm_cap > 0;
        sl.sv->need_w = false;
        const int r = sl.sv->setup();
        if (r != EDLGT_OK) {
          sl.sv.reset();
          return r;
        }
        pick = (int)i;
      }
    if (pick < 0) return EDLGT_OK;
    Slot& sl = slots[pick];
    for (int t = 0; t < 16; ++t) sl.coef[t] = t < h.n_terms ? coef[t] : 0.0;
    sl.sv->seed = seed;
    sl.sv->mv64 = sl.sv->mv32 = 0;
    const int64_t N = h.n_rows;
    const int e = warm_cap > 0 ? nearest_warm(sl.coef) : -1;
    // A warm start block approximates every wanted eigenvector, so the
    // convergence test runs at every step; a random start block is tested only
    // when the basis is full (a small Krylov space can hold converged Ritz
    // pairs while a lower or degenerate eigenvector is still missing).
    sl.sv->early_every = e >= 0 ? 1 : 0;
    if (e >= 0) {
      double w = 0.0;
      const int e2 = (k == b && k <= 8) ? bracket_warm(sl.coef, e, w) : -1;
      bool done = false;
      if (e2 >= 0) EDLGT_TRY(interpolate_warm(*sl.sv, e, e2, w, done));
      if (!done)
        EDLGT_HIP(hipMemcpyAsync(sl.sv->col(0), warm.as<S>() + (size_t)e * k * N, (size_t)b * N * sizeof(S),
                                 hipMemcpyDeviceToDevice, st));
    } else {
      EDLGT_TRY(sl.sv->randomize(sl.sv->col(0), (int64_t)b * N));
    }
    EDLGT_TRY(sl.sv->begin(k, tol, max_mv));
    sl.req = req;
    sl.active = true;
    *slot_out = pick;
    return EDLGT_OK;
  }

  // Keeps the Ritz vectors of a finished request (Tmp[:, 0:k]); when the cache
  // is full the entry nearest in coefficients is replaced, which keeps the
  // cached couplings spread over the requested range.
  int remember(Slot& sl) {
    if (warm_cap <= 0) return EDLGT_OK;
    const int64_t N = h.n_rows;
    if (!warm.p) EDLGT_TRY(warm.alloc((size_t)warm_cap * k * (size_t)h.n_rows * sizeof(S)));
    int e;
    if (warm_n < warm_cap) {
      e = warm_n++;
      warm_coef.emplace_back(sl.coef, sl.coef + 16);
    } else {
      e = nearest_warm(sl.coef);
      warm_coef[e].assign(sl.coef, sl.coef + 16);
    }
    EDLGT_HIP(hipMemcpyAsync(warm.as<S>() + (size_t)e * k * N, sl.sv->Tmp.p, (size_t)k * N * sizeof(S),
                             hipMemcpyDeviceToDevice, st));
    return EDLGT_OK;
  }

  // Fused SpMM over the active blocks of all active slots, in chunks of at most
  // 16 columns; each slot's product lands in V[:, nv:nv+b] (out_block).
  int fused_matvec() {
    auto* px = hxp.as<const void*>();
    auto* py = hyp.as<void*>();
    auto* pc = hcp.as<double>();
    int ncol = 0;
    for (auto& sl : slots) {
      if (!sl.active) continue;
      for (int c = 0; c < b; ++c) {
        px[ncol] = sl.sv->active_block() + (int64_t)c * h.n_rows;
        py[ncol] = sl.sv->out_block() + (int64_t)c * h.n_rows;
        for (int t = 0; t < 16; ++t) pc[(size_t)ncol * 16 + t] = sl.coef[t];
        sl.sv->mv64 += 1;
        ++ncol;
      }
    }
    if (ncol == 0) return EDLGT_OK;
    const size_t need = coef_pack_bytes(h, vec_type, kCoefMaxRhs);
    if (pack.bytes < need) {  // also after the matrix was coded
      drop_graphs();
      EDLGT_TRY(pack.alloc(need));
    }
    if (!use_graphs) {
      EDLGT_TRY(issue_matvec(ncol));
    } else {
      hipGraphExec_t exec = nullptr;
      for (auto& g : graphs)
        if (g.ncol == ncol && g.pack == pack.p) exec = g.exec;
      if (!exec) {
        EDLGT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int s1 = issue_matvec(ncol);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (s1 != EDLGT_OK) {
          if (graph) (void)hipGraphDestroy(graph);
          return s1;
        }
        EDLGT_HIP(e);
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        EDLGT_HIP(ei);
        graphs.push_back(StepGraph{ncol, pack.p, exec});
        ++graph_captures;
      }
      EDLGT_HIP(hipGraphLaunch(exec, st));
      ++graph_launches;
    }
    EDLGT_HIP(hipStreamSynchronize(st));  // the staging buffers are rewritten next step
    return EDLGT_OK;
  }

  // The launches of one fused product over ncol columns (eager, or recorded
  // into a graph): pointer and coefficient uploads from the pinned staging
  // buffers, then one coefficient SpMM per chunk of at most 16 columns.
  int issue_matvec(int ncol) {
    EDLGT_HIP(hipMemcpyAsync(xs.p, hxp.p, ncol * sizeof(void*), hipMemcpyHostToDevice, st));
    EDLGT_HIP(hipMemcpyAsync(ys.p, hyp.p, ncol * sizeof(void*), hipMemcpyHostToDevice, st));
    EDLGT_HIP(hipMemcpyAsync(cf.p, hcp.p, (size_t)(ncol + kCoefMaxRhs) * 16 * sizeof(double),
                             hipMemcpyHostToDevice, st));
    for (int done = 0; done < ncol; done += kCoefMaxRhs) {
      const int nb = std::min(ncol - done, kCoefMaxRhs);
      EDLGT_TRY(csr_spmm_coef(h, vec_type, nb, xs.as<void*>() + done, ys.as<void*>() + done, nullptr, nullptr,
                              cf.as<double>() + (size_t)done * 16, pack.p, st));
    }
    return EDLGT_OK;
  }

  struct Done {
    int64_t req;
    std::vector<double> theta, est;
    int64_t matvecs;
    bool converged;
  };

  int step(std::vector<Done>& out) {
    if (active() == 0) return EDLGT_OK;
    EDLGT_TRY(fused_matvec());
    for (auto& sl : slots) {
      if (!sl.active) continue;
      bool finished = false, conv = false;
      std::vector<double> th, es;
      EDLGT_TRY(sl.sv->after_matvec(finished, th, es, conv));
      if (finished) {
        if (conv) EDLGT_TRY(remember(sl));
        out.push_back(Done{sl.req, th, es, sl.sv->mv64, conv});
        sl.active = false;
        sl.req = -1;
      }
    }
    return EDLGT_OK;
  }
};

struct ServerAny {
  int vec_type;
  std::unique_ptr<Server<double>> r;
  std::unique_ptr<Server<double2>> c;
};

}  // namespace

int krylov_run(Csr& h, const Basis* basis, int precision, void* psi, double dt, int64_t n_steps,
               int m, int reorth, int d_max, double* out_marg, double* out_err, void* history,
               hipStream_t st) {
  return krylov_typed(h, basis, precision, psi, dt, n_steps, m, reorth, d_max, out_marg, out_err,
                      history, st);
}

int lanczos_run(Csr& h, int n_eigs, double tol, int max_basis, int64_t max_iter, int precision,
                int vec_type, const void* v0, double* evals, void* evecs, double* resid,
                int64_t* matvecs, hipStream_t st) {
  if (vec_type == EDLGT_VALUE_C128)
    return lanczos_typed<double2>(h, n_eigs, tol, max_basis, max_iter, precision, v0, evals,
                                  evecs, resid, matvecs, st);
  if (vec_type == EDLGT_VALUE_F64) {
    if (h.value_type != EDLGT_VALUE_F64)
      return set_error(EDLGT_ERR_INVALID_ARG, "real vectors need a real matrix");
    return lanczos_typed<double>(h, n_eigs, tol, max_basis, max_iter, precision, v0, evals,
                                 evecs, resid, matvecs, st);
  }
  return set_error(EDLGT_ERR_INVALID_ARG, "vec_type must be C128 or F64");
}

void* server_create(Csr& h, int vec_type, int max_slots, int n_eigs, double tol, int max_basis,
                    int64_t max_mv, int warm_cap, hipStream_t st, int* status) {
  auto* sa = new ServerAny();
  sa->vec_type = vec_type;
  int s;
  if (vec_type == EDLGT_VALUE_F64) {
    sa->r.reset(new Server<double>(h, st, vec_type));
    s = sa->r->setup(max_slots, n_eigs, tol, max_basis, max_mv, warm_cap);
  } else {
    sa->c.reset(new Server<double2>(h, st, vec_type));
    s = sa->c->setup(max_slots, n_eigs, tol, max_basis, max_mv, warm_cap);
  }
  *status = s;
  if (s != EDLGT_OK) {
    delete sa;
    return nullptr;
  }
  return sa;
}

void server_destroy(void* p) { delete static_cast<ServerAny*>(p); }

int server_submit(void* p, int64_t req, const double* coef, uint64_t seed, int* slot) {
  auto* sa = static_cast<ServerAny*>(p);
  return sa->r ? sa->r->submit(req, coef, seed, slot) : sa->c->submit(req, coef, seed, slot);
}

void server_trim(void* p) {
  auto* sa = static_cast<ServerAny*>(p);
  if (sa->r)
    sa->r->trim();
  else
    sa->c->trim();
}

void server_info(void* p, int* active, int* allocated, int64_t* slot_bytes, int64_t* warm_bytes) {
  auto* sa = static_cast<ServerAny*>(p);
  if (sa->r) {
    *active = sa->r->active();
    *allocated = sa->r->allocated();
    *slot_bytes = (int64_t)sa->r->slot_bytes();
    *warm_bytes = (int64_t)sa->r->shared_bytes();
  } else {
    *active = sa->c->active();
    *allocated = sa->c->allocated();
    *slot_bytes = (int64_t)sa->c->slot_bytes();
    *warm_bytes = (int64_t)sa->c->shared_bytes();
  }
}

int server_step(void* p, int max_done, int64_t* ids, double* evals, double* est, int64_t* mvs,
                int32_t* conv, int* n_done) {
  auto* sa = static_cast<ServerAny*>(p);
  *n_done = 0;
  auto emit = [&](auto& done) {
    for (auto& d : done) {
      if (*n_done >= max_done) return set_error(EDLGT_ERR_INVALID_ARG, "max_done too small");
      const int i = (*n_done)++;
      ids[i] = d.req;
      const int k = (int)d.theta.size();
      for (int j = 0; j < k; ++j) {
        evals[(int64_t)i * k + j] = d.theta[j];
        if (est) est[(int64_t)i * k + j] = d.est[j];
      }
      if (mvs) mvs[i] = d.matvecs;
      if (conv) conv[i] = d.converged ? 1 : 0;
    }
    return (int)EDLGT_OK;
  };
  if (sa->r) {
    std::vector<Server<double>::Done> done;
    EDLGT_TRY(sa->r->step(done));
    return emit(done);
  }
  std::vector<Server<double2>::Done> done;
  EDLGT_TRY(sa->c->step(done));
  return emit(done);
}

}  // namespace edlgt
