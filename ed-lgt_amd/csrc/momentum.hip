// Momentum sectors: the Bloch-sum projector B of a translation group acting on
// the sorted sector basis, the projected Hamiltonian B^H H B, and B / B^H on
// vectors.
//
// Follows edlgt/symmetries/translational_sym.py (get_momentum_basis,
// :1307-1359, and the zero-k / finite-k constructors :698-950): block
// translations are site permutations in that module's row-major coordinates
// (last axis fastest, :131-157, :508-565); the translation table holds the
// basis index of every translated configuration (:568-618); each orbit is
// represented by its smallest basis index with its axis periods p_d, and is
// kept iff (k_d p_d) % R_d == 0 on every axis (:621-695); its column is the
// normalised sum over the period box of exp(-2 pi i k.t / R) T(t)|ref>, equal
// images accumulated in box order, amplitudes |a| <= 1e-14 dropped for k != 0
// (a column that cancels completely stays, empty), rows sorted.
//
// Device: the translation table (one lookup per configuration and shift), the
// orbit scan of standard translations (every state reads its own table row:
// representative = row minimum, rank in the orbit, Bloch sum over the shifts
// that reach it; same columns as the box sum, since the phases are periodic
// on the period lattice) and the projected matrix.  Host: the orbit scan of
// the TC translations of the 1D chain (sign per representative).
//
// Projected matrix, one 64-lane wave per k-row c: for every member r of
// column c (amplitude b_r) and every stored entry (j, h) of H's row r whose
// configuration j lies in a kept column c_j with amplitude b_j, the wave emits
// (c_j, conj(b_r) h b_j) into LDS in (member, entry) order; a bitonic sort by
// (column, slot) and a sum of equal columns in slot order give the row
// (count pass -> scan -> fill pass, as the Hamiltonian assembly).
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace edlgt {
namespace {

constexpr double kTolZero = 1e-14;   // translational_sym.py:870
constexpr double kTolColNorm = 1e-30;

template <class K>
__device__ inline int64_t lower_bound_k(const K* __restrict__ a, int64_t n, K k) {
  int64_t lo = 0, len = n;
  while (len > 0) {
    const int64_t half = len >> 1;
    const int64_t mid = lo + half;
    if (a[mid] < k) {
      lo = mid + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo;
}

// trans[r * S + s] = basis index of shift s applied to configuration r
// (rolled[dest] = cfg[src[s * n_sites + dest]]); a translated configuration
// outside the sector sets *err.
template <class KeyT>
__global__ void translate_rows(const uint8_t* __restrict__ configs, int cfg_stride, int n_sites,
                               const KeyT* __restrict__ keys, int64_t n, const KeyT* __restrict__ strides,
                               const int32_t* __restrict__ src, int S, int32_t* __restrict__ trans,
                               int32_t* __restrict__ err) {
  const int64_t total = n * S;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / S;
    const int s = (int)(i % S);
    const uint8_t* cfg = configs + r * cfg_stride;
    const int32_t* sp = src + (int64_t)s * n_sites;
    KeyT key = 0;
    for (int d = 0; d < n_sites; ++d) key += (KeyT)cfg[sp[d]] * strides[d];
    const int64_t j = lower_bound_k<KeyT>(keys, n, key);
    if (j >= n || keys[j] != key) {
      atomicOr(err, 1);
      trans[i] = 0;
    } else {
      trans[i] = (int32_t)j;
    }
  }
}


// trans[r * L + t] = index of X^t configuration r, X = shift by one site after
// the local map C: X^t puts C^t(cfg[j]) on site (j + t) % L, C applied for odd
// t only (C is an involution); sign[r] = prod_j c_sign[cfg[j]].
template <class KeyT>
__global__ void translate_rows_tc(const uint8_t* __restrict__ configs, int cfg_stride, int L,
                                  const KeyT* __restrict__ keys, int64_t n, const KeyT* __restrict__ strides,
                                  const int32_t* __restrict__ c_map, const double* __restrict__ c_sign,
                                  int32_t* __restrict__ trans, double* __restrict__ sign, int32_t* __restrict__ err) {
  const int64_t total = n * L;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / L;
    const int t = (int)(i % L);
    const uint8_t* cfg = configs + r * cfg_stride;
    KeyT key = 0;
    double sg = 1.0;
    for (int j = 0; j < L; ++j) {
      const int a = cfg[j];
      const int b = (t & 1) ? c_map[a] : a;
      key += (KeyT)b * strides[(j + t) % L];
      sg *= c_sign[a];
    }
    if (t == 0) sign[r] = sg;
    const int64_t q = lower_bound_k<KeyT>(keys, n, key);
    if (q >= n || keys[q] != key) {
      atomicOr(err, 1);
      trans[i] = 0;
    } else {
      trans[i] = (int32_t)q;
    }
  }
}


// Signed image of every configuration under a site permutation with per-site
// local maps: the image puts loc_perm[j][c_j] on site site_perm[j] and carries
// the sign prod_j loc_sign[j][c_j]; out_idx[r] = its basis index (-1 and *err
// set when the image is not in the sector).  The spatial inversion of
// SU2_Model.get_parity_inversion_operator (inversion_sym.py:77-161).
template <class KeyT>
__global__ void signed_image_rows(const uint8_t* __restrict__ configs, int cfg_stride, int n_sites,
                                  const KeyT* __restrict__ keys, int64_t n, const KeyT* __restrict__ strides,
                                  const int32_t* __restrict__ site_perm, const int32_t* __restrict__ loc_perm,
                                  const double* __restrict__ loc_sign, int d_max, int64_t* __restrict__ out_idx,
                                  double* __restrict__ out_sign, int32_t* __restrict__ err) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const uint8_t* cfg = configs + r * cfg_stride;
    KeyT key = 0;
    double sg = 1.0;
    for (int j = 0; j < n_sites; ++j) {
      const int a = cfg[j];
      key += (KeyT)loc_perm[j * d_max + a] * strides[site_perm[j]];
      sg *= loc_sign[j * d_max + a];
    }
    const int64_t q = lower_bound_k<KeyT>(keys, n, key);
    if (q >= n || keys[q] != key) {
      atomicOr(err, 1);
      out_idx[r] = -1;
    } else {
      out_idx[r] = q;
    }
    out_sign[r] = sg;
  }
}

__device__ inline double2 cm(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 conj2(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline double2 val_c(const double* v, int64_t p) { return make_double2(v[p], 0.0); }
__device__ inline double2 val_c(const double2* v, int64_t p) { return v[p]; }

__device__ inline int wave_excl(int v, int* total) {
  const int lane = threadIdx.x & 63;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  *total = __shfl(x, 63, 64);
  return x - v;
}

template <int kCap, bool kFill, class VT>
__global__ void __launch_bounds__(64) momentum_rows(
    int64_t n_cols, const int64_t* __restrict__ bptr, const int32_t* __restrict__ brow,
    const double2* __restrict__ bval, const int32_t* __restrict__ row_col, const double2* __restrict__ row_val,
    const int64_t* __restrict__ hrp, const int32_t* __restrict__ hcol, const VT* __restrict__ hval,
    int32_t* __restrict__ count, const int64_t* __restrict__ orp, int32_t* __restrict__ ocol,
    double* __restrict__ oval, int out_real, int32_t* __restrict__ err) {
  __shared__ int32_t scol[kCap];
  __shared__ uint16_t sidx[kCap];
  __shared__ double2 sval[kFill ? kCap : 1];
  const int lane = threadIdx.x;
  for (int64_t c = blockIdx.x; c < n_cols; c += gridDim.x) {
    int T = 0;
    bool overflow = false;
    for (int64_t m = bptr[c]; m < bptr[c + 1] && !overflow; ++m) {
      const int64_t r = brow[m];
      const double2 a = conj2(bval[m]);
      for (int64_t pb = hrp[r]; pb < hrp[r + 1]; pb += 64) {
        const int64_t p = pb + lane;
        int cj = -1;
        double2 v = make_double2(0.0, 0.0);
        if (p < hrp[r + 1]) {
          const int32_t j = hcol[p];
          cj = row_col[j];
          if (kFill && cj >= 0) v = cm(cm(a, val_c(hval, p)), row_val[j]);
        }
        int tot;
        const int off = wave_excl(cj >= 0 ? 1 : 0, &tot);
        if (T + tot > kCap) {
          overflow = true;
          break;
        }
        if (cj >= 0) {
          scol[T + off] = cj;
          sidx[T + off] = (uint16_t)(T + off);
          if (kFill) sval[T + off] = v;
        }
        T += tot;
      }
    }
    if (overflow) {
      if (lane == 0) atomicOr(err, 2);
      if (!kFill) count[c] = 0;
      __syncthreads();
      continue;
    }
    int P2 = 1;
    while (P2 < T) P2 <<= 1;
    for (int i = T + lane; i < P2; i += 64) {
      scol[i] = 0x7fffffff;
      sidx[i] = 0xffff;
    }
    __syncthreads();
    for (int k = 2; k <= P2; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = lane; i < P2; i += 64) {
          const int ixj = i ^ j;
          if (ixj > i) {
            const int32_t ka = scol[i], kb = scol[ixj];
            const uint16_t ia = sidx[i], ib = sidx[ixj];
            const bool a_gt_b = ka > kb || (ka == kb && ia > ib);
            const bool up = (i & k) == 0;
            if (a_gt_b == up) {
              scol[i] = kb; scol[ixj] = ka;
              sidx[i] = ib; sidx[ixj] = ia;
            }
          }
        }
        __syncthreads();
      }
    }
    int rank_base = 0;
    for (int ib = 0; ib < T; ib += 64) {
      const int i = ib + lane;
      const bool head = i < T && (i == 0 || scol[i] != scol[i - 1]);
      int tot;
      const int rk = wave_excl(head ? 1 : 0, &tot);
      if (kFill && head) {
        double2 v = sval[sidx[i]];
        for (int q = i + 1; q < T && scol[q] == scol[i]; ++q) {
          const double2 w = sval[sidx[q]];
          v.x += w.x;
          v.y += w.y;
        }
        const int64_t o = orp[c] + rank_base + rk;
        ocol[o] = scol[i];
        if (out_real) {
          oval[o] = v.x;
        } else {
          oval[2 * o] = v.x;
          oval[2 * o + 1] = v.y;
        }
      }
      rank_base += tot;
    }
    if (!kFill && lane == 0) count[c] = rank_base;
    __syncthreads();
  }
}

template <bool kForward>
__global__ void apply_b(int64_t n_rows, int64_t n_cols, const int64_t* __restrict__ bptr,
                        const int32_t* __restrict__ brow, const double2* __restrict__ bval,
                        const int32_t* __restrict__ row_col, const double2* __restrict__ row_val,
                        const double2* __restrict__ in, double2* __restrict__ out) {
  if (kForward) {  // out[r] = b_r in[c_r]   (psi = B psi_k)
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_rows;
         r += (int64_t)gridDim.x * blockDim.x) {
      const int32_t c = row_col[r];
      out[r] = c >= 0 ? cm(row_val[r], in[c]) : make_double2(0.0, 0.0);
    }
  } else {  // out[c] = sum_members conj(b_r) in[r]   (psi_k = B^H psi)
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n_cols;
         c += (int64_t)gridDim.x * blockDim.x) {
      double2 s = make_double2(0.0, 0.0);
      for (int64_t m = bptr[c]; m < bptr[c + 1]; ++m) {
        const double2 t = cm(conj2(bval[m]), in[brow[m]]);
        s.x += t.x;
        s.y += t.y;
      }
      out[c] = s;
    }
  }
}

// Translation group geometry for the device orbit scan: shift s is decoded
// row-major over the axis counts R (s_d = (s / sstride_d) % R_d), kd = k mod R.
struct OrbitGeo {
  int n_dims, S, max_r;
  int r[3], sstride[3], kd[3];
};

__device__ inline int32_t row_min(const int32_t* __restrict__ row, int S) {
  int32_t mn = row[0];
  for (int s = 1; s < S; ++s) mn = min(mn, row[s]);
  return mn;
}

// exp(-2 pi i k.s / R) == 1 for shift s, tested exactly on the integers
// sum_d kd_d s_d (S / R_d) mod S
__device__ inline bool k_trivial(const OrbitGeo& g, int s) {
  int64_t num = 0;
  for (int d = 0; d < g.n_dims; ++d)
    num += (int64_t)g.kd[d] * ((s / g.sstride[d]) % g.r[d]) * (g.S / g.r[d]);
  return num % g.S == 0;
}

// Per state j (one thread, its translation-table row): whether j is the
// smallest index of its orbit and the orbit is kept ((k_d p_d) % R_d == 0 with
// p_d the axis periods), and the stored entries of its column: the orbit size
// S / |H| (H = shifts fixing j), or 0 when k is not trivial on H (then every
// Bloch sum of the orbit cancels and the column stays, empty).
__global__ void orbit_scan(const int32_t* __restrict__ trans, int64_t n, OrbitGeo g, int32_t* __restrict__ keep_rep,
                           int32_t* __restrict__ n_members) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int32_t* row = trans + j * g.S;
    int32_t flag = 0, cnt = 0;
    if (row_min(row, g.S) == (int32_t)j) {
      bool keep = true;
      for (int d = 0; d < g.n_dims; ++d) {
        int p = g.r[d];
        for (int q = 1; q <= g.r[d]; ++q)
          if (row[(q % g.r[d]) * g.sstride[d]] == (int32_t)j) {
            p = q;
            break;
          }
        if ((g.kd[d] * p) % g.r[d] != 0) keep = false;
      }
      if (keep) {
        flag = 1;
        int h = 0;
        bool triv = true;
        for (int s = 0; s < g.S; ++s)
          if (row[s] == (int32_t)j) {
            ++h;
            triv = triv && k_trivial(g, s);
          }
        cnt = triv ? g.S / h : 0;
      }
    }
    keep_rep[j] = flag;
    n_members[j] = cnt;
  }
}

// Per state j with representative m (its row's minimum): the column of m
// (col_of = exclusive scan of keep_rep), the rank of j in the orbit
// (#{s : T_s j < j} / |H|: the row lists every member |H| times) and the
// amplitude sum over {s : T_s j = m} of exp(+2 pi i k.s / R) (the shifts
// t = -s carrying m to j), normalised by sqrt(S |H|) (S / |H| members of
// modulus |H|).  ptab[d][q] = exp(-2 pi i kd_d q / R_d).
__global__ void orbit_fill(const int32_t* __restrict__ trans, int64_t n, OrbitGeo g, const double2* __restrict__ ptab,
                           const int32_t* __restrict__ keep_rep, const int32_t* __restrict__ n_members,
                           const int64_t* __restrict__ col_of, const int64_t* __restrict__ first,
                           int64_t* __restrict__ col_ptr, int32_t* __restrict__ row_idx, double2* __restrict__ data,
                           int32_t* __restrict__ row_col, double2* __restrict__ row_val) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int32_t* row = trans + j * g.S;
    const int32_t m = row_min(row, g.S);
    int32_t c = -1;
    double2 v = make_double2(0.0, 0.0);
    if (keep_rep[m]) {
      const int64_t col = col_of[m];
      if (m == (int32_t)j) col_ptr[col] = first[m];
      if (n_members[m] > 0) {
        int h = 0, less = 0;
        double2 amp = make_double2(0.0, 0.0);
        for (int s = 0; s < g.S; ++s) {
          const int32_t x = row[s];
          h += x == (int32_t)j;
          less += x < (int32_t)j;
          if (x == m) {
            double2 ph = make_double2(1.0, 0.0);
            for (int d = 0; d < g.n_dims; ++d) {
              const int sd = (s / g.sstride[d]) % g.r[d];
              ph = cm(ph, ptab[d * g.max_r + (g.r[d] - sd) % g.r[d]]);
            }
            amp.x += ph.x;
            amp.y += ph.y;
          }
        }
        const double inv = 1.0 / sqrt((double)g.S * (double)h);
        v = make_double2(amp.x * inv, amp.y * inv);
        const int64_t o = first[m] + less / h;
        row_idx[o] = (int32_t)j;
        data[o] = v;
        c = (int32_t)col;
      }
    }
    row_col[j] = c;
    row_val[j] = v;
  }
}

int check_flag(const DevBuf& err, hipStream_t st, int32_t* out) {
  EDLGT_HIP(hipMemcpyAsync(out, err.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  return EDLGT_OK;
}

}  // namespace

// Host orbit scan of the TC translations of a 1D chain of L sites over the
// table trans[N][L] (shift t = X^t): one column per orbit whose smallest index
// is i and whose period p satisfies (k p) % L == 0, the p images summed with
// the phases exp(-2 pi i k t / L) (evaluated directly) times tc_sign[i] (the C
// sign of configuration i) on odd t, equal images accumulated in shift order;
// a k = 0 column whose norm^2 is within 1e-8 of zero stays empty.
static int orbit_columns_tc(const std::vector<int32_t>& trans, int64_t N, int64_t L, int32_t k,
                            const double* tc_sign, hipStream_t st, Momentum* out) {
  const int64_t kd = ((k % L) + L) % L;
  const bool finite = k != 0;
  using cd = std::complex<double>;
  std::vector<int64_t> col_ptr(1, 0);
  std::vector<int32_t> row_idx;
  std::vector<double2> data;
  std::vector<int32_t> row_col((size_t)N, -1);
  std::vector<double2> row_val((size_t)N, make_double2(0.0, 0.0));
  std::vector<int32_t> img;
  std::vector<cd> amp;
  for (int64_t i = 0; i < N; ++i) {
    const int32_t* row = trans.data() + (size_t)i * L;
    int32_t mn = row[0];
    for (int64_t s = 1; s < L; ++s) mn = std::min(mn, row[s]);
    if (mn != (int32_t)i) continue;  // not the orbit's smallest index
    int64_t p = L;
    for (int64_t q = 1; q <= L; ++q)
      if (row[q % L] == (int32_t)i) {
        p = q;
        break;
      }
    if ((k * p) % L != 0) continue;
    img.clear();
    amp.clear();
    for (int64_t flat = 0; flat < p; ++flat) {
      cd ph(1.0, 0.0);
      if (finite) ph = std::exp(cd(0.0, -1.0) * (2.0 * M_PI * ((double)(kd * flat) / (double)L)));
      if (flat & 1) ph *= tc_sign[i];
      const int32_t im = row[flat];
      size_t u = 0;
      while (u < img.size() && img[u] != im) ++u;
      if (u == img.size()) {
        img.push_back(im);
        amp.push_back(ph);
      } else {
        amp[u] += ph;
      }
    }
    double nrm = 0.0;
    size_t kept = 0;
    for (size_t u = 0; u < img.size(); ++u) {
      if (finite && std::abs(amp[u]) <= kTolZero) continue;
      img[kept] = img[u];
      amp[kept] = amp[u];
      nrm += std::norm(amp[u]);
      ++kept;
    }
    img.resize(kept);
    amp.resize(kept);
    const int32_t col = (int32_t)(col_ptr.size() - 1);
    const bool empty = finite ? (nrm <= kTolColNorm || kept == 0) : (std::fabs(nrm) <= 1e-8);
    if (!empty) {
      std::vector<size_t> ord(kept);
      for (size_t u = 0; u < kept; ++u) ord[u] = u;
      std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return img[a] < img[b]; });
      const double inv = 1.0 / std::sqrt(nrm);
      for (size_t u : ord) {
        const cd v = amp[u] * inv;
        row_idx.push_back(img[u]);
        data.push_back(make_double2(v.real(), v.imag()));
        row_col[(size_t)img[u]] = col;
        row_val[(size_t)img[u]] = make_double2(v.real(), v.imag());
      }
    }
    col_ptr.push_back((int64_t)row_idx.size());
  }
  out->n_rows = N;
  out->n_cols = (int64_t)col_ptr.size() - 1;
  out->nnz = (int64_t)row_idx.size();
  out->is_complex = finite ? 1 : 0;
  EDLGT_TRY(upload(out->col_ptr, col_ptr.data(), col_ptr.size(), st));
  const int32_t z32 = 0;
  const double2 z2 = make_double2(0.0, 0.0);
  EDLGT_TRY(upload(out->row_idx, row_idx.empty() ? &z32 : row_idx.data(), std::max<size_t>(row_idx.size(), 1), st));
  EDLGT_TRY(upload(out->data, data.empty() ? &z2 : data.data(), std::max<size_t>(data.size(), 1), st));
  EDLGT_TRY(upload(out->row_col, row_col.empty() ? &z32 : row_col.data(), std::max<size_t>(row_col.size(), 1), st));
  EDLGT_TRY(upload(out->row_val, row_val.empty() ? &z2 : row_val.data(), std::max<size_t>(row_val.size(), 1), st));
  EDLGT_HIP(hipStreamSynchronize(st));
  return EDLGT_OK;
}

// Device orbit scan over the device translation table (standard
// translations): the same columns as the reference's period-box sum, each state placed by its
// own table row (representative, rank, Bloch sum), so the table never leaves
// the device.  Two passes: per-state flags and column sizes, scanned into
// column numbers and column starts, then every state writes its entry.
static int orbit_columns_device(const DevBuf& trans, int64_t N, int64_t S, const std::vector<int64_t>& R,
                                const int32_t* k_vals, hipStream_t st, Momentum* out) {
  OrbitGeo g{};
  g.n_dims = (int)R.size();
  g.S = (int)S;
  g.max_r = (int)*std::max_element(R.begin(), R.end());
  bool finite = false;
  for (int d = g.n_dims - 1, stride = 1; d >= 0; --d) {
    g.r[d] = (int)R[d];
    g.sstride[d] = stride;
    stride *= g.r[d];
    g.kd[d] = (int)(((k_vals[d] % R[d]) + R[d]) % R[d]);
    finite = finite || k_vals[d] != 0;
  }
  // per-axis phase tables by repeated multiplication (the reference's phase table)
  using cd = std::complex<double>;
  std::vector<double2> ptab((size_t)g.n_dims * g.max_r, make_double2(1.0, 0.0));
  for (int d = 0; d < g.n_dims; ++d) {
    const cd step = std::exp(cd(0.0, -2.0 * M_PI * (double)g.kd[d] / (double)g.r[d]));
    cd v(1.0, 0.0);
    for (int q = 1; q < g.r[d]; ++q) {
      v *= step;
      ptab[(size_t)d * g.max_r + q] = make_double2(v.real(), v.imag());
    }
  }
  DevBuf b_ptab, b_keep, b_cnt, b_col, b_first;
  EDLGT_TRY(upload(b_ptab, ptab.data(), ptab.size(), st));
  EDLGT_TRY(b_keep.alloc((size_t)std::max<int64_t>(N, 1) * sizeof(int32_t)));
  EDLGT_TRY(b_cnt.alloc((size_t)std::max<int64_t>(N, 1) * sizeof(int32_t)));
  EDLGT_TRY(b_col.alloc((size_t)(N + 1) * sizeof(int64_t)));
  EDLGT_TRY(b_first.alloc((size_t)(N + 1) * sizeof(int64_t)));
  const unsigned grid = grid_for(N, 256, 65536);
  if (N > 0) {
    hipLaunchKernelGGL(orbit_scan, dim3(grid), dim3(256), 0, st, trans.as<int32_t>(), N, g, b_keep.as<int32_t>(),
                       b_cnt.as<int32_t>());
    EDLGT_HIP(hipGetLastError());
  }
  EDLGT_TRY(exclusive_scan_i32_i64(b_keep.as<int32_t>(), b_col.as<int64_t>(), N, st));
  EDLGT_TRY(exclusive_scan_i32_i64(b_cnt.as<int32_t>(), b_first.as<int64_t>(), N, st));
  int64_t tot[2] = {0, 0};
  EDLGT_HIP(hipMemcpyAsync(&tot[0], b_col.as<int64_t>() + N, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipMemcpyAsync(&tot[1], b_first.as<int64_t>() + N, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  const int64_t n_cols = tot[0], nnz = tot[1];
  EDLGT_TRY(out->col_ptr.alloc((size_t)(n_cols + 1) * sizeof(int64_t)));
  EDLGT_TRY(out->row_idx.alloc((size_t)std::max<int64_t>(nnz, 1) * sizeof(int32_t)));
  EDLGT_TRY(out->data.alloc((size_t)std::max<int64_t>(nnz, 1) * sizeof(double2)));
  EDLGT_TRY(out->row_col.alloc((size_t)std::max<int64_t>(N, 1) * sizeof(int32_t)));
  EDLGT_TRY(out->row_val.alloc((size_t)std::max<int64_t>(N, 1) * sizeof(double2)));
  EDLGT_HIP(hipMemcpyAsync(out->col_ptr.as<int64_t>() + n_cols, &tot[1], sizeof(int64_t), hipMemcpyHostToDevice, st));
  if (N > 0) {
    hipLaunchKernelGGL(orbit_fill, dim3(grid), dim3(256), 0, st, trans.as<int32_t>(), N, g, b_ptab.as<double2>(),
                       b_keep.as<int32_t>(), b_cnt.as<int32_t>(), b_col.as<int64_t>(), b_first.as<int64_t>(),
                       out->col_ptr.as<int64_t>(), out->row_idx.as<int32_t>(), out->data.as<double2>(),
                       out->row_col.as<int32_t>(), out->row_val.as<double2>());
    EDLGT_HIP(hipGetLastError());
  }
  EDLGT_HIP(hipStreamSynchronize(st));
  out->n_rows = N;
  out->n_cols = n_cols;
  out->nnz = nnz;
  out->is_complex = finite ? 1 : 0;
  return EDLGT_OK;
}

int momentum_build(const Basis& B, int n_dims, const int32_t* lvals, const int32_t* unit_cell,
                   const int32_t* k_vals, hipStream_t st, Momentum* out) {
  if (n_dims < 1 || n_dims > 3) return set_error(EDLGT_ERR_INVALID_ARG, "n_dims must be 1..3");
  int64_t n_sites = 1, S = 1;
  std::vector<int64_t> R(n_dims);
  for (int d = 0; d < n_dims; ++d) {
    if (lvals[d] < 1 || unit_cell[d] < 1 || lvals[d] % unit_cell[d] != 0)
      return set_error(EDLGT_ERR_INVALID_ARG, "unit_cell must divide lvals on every axis");
    R[d] = lvals[d] / unit_cell[d];
    n_sites *= lvals[d];
    S *= R[d];
  }
  if (n_sites != B.n_sites) return set_error(EDLGT_ERR_INVALID_ARG, "prod(lvals) differs from the basis sites");
  if (S > 4096) return set_error(EDLGT_ERR_INVALID_ARG, "more than 4096 translations");
  const int64_t N = B.n_states;
  if (N > 0x7fffffffLL) return set_error(EDLGT_ERR_INDEX_OVERFLOW, "more than 2^31 - 1 sector states");
  if (N * S > (int64_t)1 << 36) return set_error(EDLGT_ERR_INDEX_OVERFLOW, "translation table too large");
  // site permutation of every block shift (row-major coordinates, last axis fastest)
  std::vector<int32_t> src((size_t)S * n_sites);
  std::vector<int64_t> t(n_dims), c(n_dims);
  for (int64_t s = 0; s < S; ++s) {
    int64_t f = s;
    for (int d = n_dims - 1; d >= 0; --d) {
      t[d] = f % R[d];
      f /= R[d];
    }
    for (int64_t site = 0; site < n_sites; ++site) {
      int64_t q = site;
      for (int d = n_dims - 1; d >= 0; --d) {
        c[d] = q % lvals[d];
        q /= lvals[d];
      }
      int64_t dest = 0;
      for (int d = 0; d < n_dims; ++d) dest = dest * lvals[d] + (c[d] + t[d] * unit_cell[d]) % lvals[d];
      src[(size_t)s * n_sites + dest] = (int32_t)site;
    }
  }
  // device translation table
  DevBuf b_src, b_str, b_trans, b_err;
  EDLGT_TRY(upload(b_src, src.data(), src.size(), st));
  if (B.key_words == 2)
    EDLGT_TRY(upload(b_str, B.strides128.data(), B.strides128.size(), st));
  else
    EDLGT_TRY(upload(b_str, B.strides.data(), B.strides.size(), st));
  EDLGT_TRY(b_trans.alloc((size_t)std::max<int64_t>(N * S, 1) * sizeof(int32_t)));
  EDLGT_TRY(b_err.alloc(sizeof(int32_t)));
  EDLGT_HIP(hipMemsetAsync(b_err.p, 0, sizeof(int32_t), st));
  const unsigned g = grid_for(N * S, 256, 65536);
  if (B.key_words == 2)
    hipLaunchKernelGGL(translate_rows<u128>, dim3(g), dim3(256), 0, st, B.configs.as<uint8_t>(), B.cfg_stride,
                       (int)n_sites, B.keys.as<u128>(), N, b_str.as<u128>(), b_src.as<int32_t>(), (int)S,
                       b_trans.as<int32_t>(), b_err.as<int32_t>());
  else
    hipLaunchKernelGGL(translate_rows<uint64_t>, dim3(g), dim3(256), 0, st, B.configs.as<uint8_t>(),
                       B.cfg_stride, (int)n_sites, B.keys.as<uint64_t>(), N, b_str.as<uint64_t>(),
                       b_src.as<int32_t>(), (int)S, b_trans.as<int32_t>(), b_err.as<int32_t>());
  EDLGT_HIP(hipGetLastError());
  int32_t flag = 0;
  EDLGT_TRY(check_flag(b_err, st, &flag));
  if (flag) return set_error(EDLGT_ERR_KEY_NOT_FOUND, "the sector is not closed under the translations");
  return orbit_columns_device(b_trans, N, S, R, k_vals, st, out);
}

int basis_signed_image(const Basis& B, const int32_t* site_perm, const int32_t* loc_perm, const double* loc_sign,
                       int d_max, int64_t* out_idx, double* out_sign, hipStream_t st) {
  const int n_sites = B.n_sites;
  const int64_t N = B.n_states;
  if (d_max < 1 || d_max > 256) return set_error(EDLGT_ERR_INVALID_ARG, "d_max must be in [1, 256]");
  std::vector<int> seen(n_sites, 0);
  for (int j = 0; j < n_sites; ++j) {
    if (site_perm[j] < 0 || site_perm[j] >= n_sites || seen[site_perm[j]]++)
      return set_error(EDLGT_ERR_INVALID_ARG, "site_perm is not a permutation of the sites");
    if (B.loc_dims[j] > d_max) return set_error(EDLGT_ERR_INVALID_ARG, "d_max below a local dimension");
  }
  for (int j = 0; j < n_sites; ++j)
    for (int a = 0; a < B.loc_dims[j]; ++a) {
      const int b = loc_perm[j * d_max + a];
      if (b < 0 || b >= B.loc_dims[site_perm[j]])
        return set_error(EDLGT_ERR_INVALID_ARG, "loc_perm maps a local state outside its image site's basis");
    }
  if (N == 0) return EDLGT_OK;
  DevBuf b_sp, b_lp, b_ls, b_str, b_idx, b_sg, b_err;
  EDLGT_TRY(upload(b_sp, site_perm, (size_t)n_sites, st));
  EDLGT_TRY(upload(b_lp, loc_perm, (size_t)n_sites * d_max, st));
  EDLGT_TRY(upload(b_ls, loc_sign, (size_t)n_sites * d_max, st));
  if (B.key_words == 2)
    EDLGT_TRY(upload(b_str, B.strides128.data(), B.strides128.size(), st));
  else
    EDLGT_TRY(upload(b_str, B.strides.data(), B.strides.size(), st));
  EDLGT_TRY(b_idx.alloc((size_t)N * sizeof(int64_t)));
  EDLGT_TRY(b_sg.alloc((size_t)N * sizeof(double)));
  EDLGT_TRY(b_err.alloc(sizeof(int32_t)));
  EDLGT_HIP(hipMemsetAsync(b_err.p, 0, sizeof(int32_t), st));
  const unsigned g = grid_for(N, 256, 65536);
  if (B.key_words == 2)
    hipLaunchKernelGGL(signed_image_rows<u128>, dim3(g), dim3(256), 0, st, B.configs.as<uint8_t>(), B.cfg_stride,
                       n_sites, B.keys.as<u128>(), N, b_str.as<u128>(), b_sp.as<int32_t>(), b_lp.as<int32_t>(),
                       b_ls.as<double>(), d_max, b_idx.as<int64_t>(), b_sg.as<double>(), b_err.as<int32_t>());
  else
    hipLaunchKernelGGL(signed_image_rows<uint64_t>, dim3(g), dim3(256), 0, st, B.configs.as<uint8_t>(),
                       B.cfg_stride, n_sites, B.keys.as<uint64_t>(), N, b_str.as<uint64_t>(), b_sp.as<int32_t>(),
                       b_lp.as<int32_t>(), b_ls.as<double>(), d_max, b_idx.as<int64_t>(), b_sg.as<double>(),
                       b_err.as<int32_t>());
  EDLGT_HIP(hipGetLastError());
  EDLGT_HIP(hipMemcpyAsync(out_idx, b_idx.p, (size_t)N * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipMemcpyAsync(out_sign, b_sg.p, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, st));
  int32_t flag = 0;
  EDLGT_TRY(check_flag(b_err, st, &flag));  // synchronises
  if (flag) return set_error(EDLGT_ERR_KEY_NOT_FOUND, "the sector is not closed under the map");
  return EDLGT_OK;
}

int momentum_build_tc(const Basis& B, int d_loc, const int32_t* c_map, const double* c_sign, int32_t k,
                      hipStream_t st, Momentum* out) {
  const int64_t N = B.n_states;
  const int L = B.n_sites;
  if (L < 1 || d_loc < 1 || d_loc > 256) return set_error(EDLGT_ERR_INVALID_ARG, "bad TC symmetry arguments");
  for (int j = 0; j < L; ++j)
    if (B.loc_dims[j] > d_loc) return set_error(EDLGT_ERR_INVALID_ARG, "local dimension above the TC map size");
  for (int a = 0; a < d_loc; ++a)
    if (c_map[a] < 0 || c_map[a] >= d_loc || c_map[c_map[a]] != a)
      return set_error(EDLGT_ERR_INVALID_ARG, "the TC local map must be an involution of [0, d_loc)");
  if (N > 0x7fffffffLL) return set_error(EDLGT_ERR_INDEX_OVERFLOW, "more than 2^31 - 1 sector states");
  DevBuf b_map, b_sgn, b_str, b_trans, b_sign, b_err;
  std::vector<int32_t> cm_(c_map, c_map + d_loc);
  std::vector<double> cs_(c_sign, c_sign + d_loc);
  EDLGT_TRY(upload(b_map, cm_.data(), cm_.size(), st));
  EDLGT_TRY(upload(b_sgn, cs_.data(), cs_.size(), st));
  if (B.key_words == 2)
    EDLGT_TRY(upload(b_str, B.strides128.data(), B.strides128.size(), st));
  else
    EDLGT_TRY(upload(b_str, B.strides.data(), B.strides.size(), st));
  EDLGT_TRY(b_trans.alloc((size_t)std::max<int64_t>(N * L, 1) * sizeof(int32_t)));
  EDLGT_TRY(b_sign.alloc((size_t)std::max<int64_t>(N, 1) * sizeof(double)));
  EDLGT_TRY(b_err.alloc(sizeof(int32_t)));
  EDLGT_HIP(hipMemsetAsync(b_err.p, 0, sizeof(int32_t), st));
  const unsigned g = grid_for(N * L, 256, 65536);
  if (B.key_words == 2)
    hipLaunchKernelGGL(translate_rows_tc<u128>, dim3(g), dim3(256), 0, st, B.configs.as<uint8_t>(), B.cfg_stride,
                       L, B.keys.as<u128>(), N, b_str.as<u128>(), b_map.as<int32_t>(), b_sgn.as<double>(),
                       b_trans.as<int32_t>(), b_sign.as<double>(), b_err.as<int32_t>());
  else
    hipLaunchKernelGGL(translate_rows_tc<uint64_t>, dim3(g), dim3(256), 0, st, B.configs.as<uint8_t>(),
                       B.cfg_stride, L, B.keys.as<uint64_t>(), N, b_str.as<uint64_t>(), b_map.as<int32_t>(),
                       b_sgn.as<double>(), b_trans.as<int32_t>(), b_sign.as<double>(), b_err.as<int32_t>());
  EDLGT_HIP(hipGetLastError());
  std::vector<int32_t> trans((size_t)N * L);
  std::vector<double> sign((size_t)N);
  if (N > 0) {
    EDLGT_HIP(hipMemcpyAsync(trans.data(), b_trans.p, trans.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    EDLGT_HIP(hipMemcpyAsync(sign.data(), b_sign.p, sign.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  int32_t flag = 0;
  EDLGT_TRY(check_flag(b_err, st, &flag));
  if (flag) return set_error(EDLGT_ERR_KEY_NOT_FOUND, "the sector is not closed under the TC translations");
  return orbit_columns_tc(trans, N, L, k, sign.data(), st, out);
}

int momentum_copy(const Momentum& m, int64_t* col_ptr, int32_t* row_idx, double* data) {
  EDLGT_HIP(hipMemcpy(col_ptr, m.col_ptr.p, (size_t)(m.n_cols + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (m.nnz > 0) {
    EDLGT_HIP(hipMemcpy(row_idx, m.row_idx.p, (size_t)m.nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
    EDLGT_HIP(hipMemcpy(data, m.data.p, (size_t)m.nnz * sizeof(double2), hipMemcpyDeviceToHost));
  }
  return EDLGT_OK;
}

int momentum_apply(const Momentum& m, int forward, const void* in, void* out, hipStream_t st) {
  const int64_t n = forward ? m.n_rows : m.n_cols;
  const unsigned g = grid_for(n, 256, 65536);
  if (forward)
    hipLaunchKernelGGL(apply_b<true>, dim3(g), dim3(256), 0, st, m.n_rows, m.n_cols, m.col_ptr.as<int64_t>(),
                       m.row_idx.as<int32_t>(), m.data.as<double2>(), m.row_col.as<int32_t>(),
                       m.row_val.as<double2>(), static_cast<const double2*>(in), static_cast<double2*>(out));
  else
    hipLaunchKernelGGL(apply_b<false>, dim3(g), dim3(256), 0, st, m.n_rows, m.n_cols, m.col_ptr.as<int64_t>(),
                       m.row_idx.as<int32_t>(), m.data.as<double2>(), m.row_col.as<int32_t>(),
                       m.row_val.as<double2>(), static_cast<const double2*>(in), static_cast<double2*>(out));
  EDLGT_HIP(hipGetLastError());
  return EDLGT_OK;
}

namespace {
template <int kCap, class VT>
int launch_mrows(bool fill, const Csr& h, const Momentum& m, const VT* hval, int32_t* count, Csr* out,
                 int out_real, int32_t* err, hipStream_t st) {
  const unsigned grid = (unsigned)std::min<int64_t>(std::max<int64_t>(m.n_cols, 1), 1 << 18);
  if (fill)
    hipLaunchKernelGGL((momentum_rows<kCap, true, VT>), dim3(grid), dim3(64), 0, st, m.n_cols,
                       m.col_ptr.as<int64_t>(), m.row_idx.as<int32_t>(), m.data.as<double2>(),
                       m.row_col.as<int32_t>(), m.row_val.as<double2>(), h.row_ptr.as<int64_t>(),
                       h.col.as<int32_t>(), hval, count, out->row_ptr.as<int64_t>(), out->col.as<int32_t>(),
                       out->val.as<double>(), out_real, err);
  else
    hipLaunchKernelGGL((momentum_rows<kCap, false, VT>), dim3(grid), dim3(64), 0, st, m.n_cols,
                       m.col_ptr.as<int64_t>(), m.row_idx.as<int32_t>(), m.data.as<double2>(),
                       m.row_col.as<int32_t>(), m.row_val.as<double2>(), h.row_ptr.as<int64_t>(),
                       h.col.as<int32_t>(), hval, count, nullptr, nullptr, nullptr, out_real, err);
  EDLGT_HIP(hipGetLastError());
  return EDLGT_OK;
}

template <class VT>
int launch_mrows_cap(int cap, bool fill, const Csr& h, const Momentum& m, const VT* hval, int32_t* count,
                     Csr* out, int out_real, int32_t* err, hipStream_t st) {
  if (cap <= 256) return launch_mrows<256, VT>(fill, h, m, hval, count, out, out_real, err, st);
  if (cap <= 512) return launch_mrows<512, VT>(fill, h, m, hval, count, out, out_real, err, st);
  if (cap <= 1024) return launch_mrows<1024, VT>(fill, h, m, hval, count, out, out_real, err, st);
  return launch_mrows<2048, VT>(fill, h, m, hval, count, out, out_real, err, st);
}
}  // namespace

int csr_momentum(const Csr& h, const Momentum& m, hipStream_t st, Csr* out) {
  if (h.n_rows != m.n_rows) return set_error(EDLGT_ERR_INVALID_ARG, "matrix and momentum basis sizes differ");
  if (h.n_terms != 1) return set_error(EDLGT_ERR_INVALID_ARG, "multi-term matrices are not projected");
  if (!h.val.p) return set_error(EDLGT_ERR_INVALID_ARG, "the matrix values were released");
  // raw entries per k-row bound: members x longest H row
  std::vector<int64_t> cp((size_t)m.n_cols + 1);
  EDLGT_HIP(hipMemcpy(cp.data(), m.col_ptr.p, cp.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
  int64_t max_members = 0;
  for (int64_t c = 0; c < m.n_cols; ++c) max_members = std::max(max_members, cp[c + 1] - cp[c]);
  std::vector<int64_t> rp((size_t)h.n_rows + 1);
  EDLGT_HIP(hipMemcpy(rp.data(), h.row_ptr.p, rp.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
  int64_t max_row = 0;
  for (int64_t r = 0; r < h.n_rows; ++r) max_row = std::max(max_row, rp[r + 1] - rp[r]);
  const int64_t bound = std::max<int64_t>(1, max_members * max_row);
  if (bound > 2048) return set_error(EDLGT_ERR_INDEX_OVERFLOW, "more than 2048 raw entries in one momentum row");
  const int cap = (int)bound;
  const bool real_out = h.value_type == EDLGT_VALUE_F64 && !m.is_complex;
  out->n_rows = m.n_cols;
  out->value_type = real_out ? EDLGT_VALUE_F64 : EDLGT_VALUE_C128;
  EDLGT_TRY(out->row_ptr.alloc((size_t)(m.n_cols + 1) * sizeof(int64_t)));
  DevBuf counts, err;
  EDLGT_TRY(counts.alloc((size_t)std::max<int64_t>(m.n_cols, 1) * sizeof(int32_t)));
  EDLGT_TRY(err.alloc(sizeof(int32_t)));
  EDLGT_HIP(hipMemsetAsync(err.p, 0, sizeof(int32_t), st));
  const bool cplx_h = h.value_type == EDLGT_VALUE_C128;
  if (m.n_cols > 0) {
    if (cplx_h)
      EDLGT_TRY(launch_mrows_cap<double2>(cap, false, h, m, h.val.as<double2>(), counts.as<int32_t>(), out,
                                          real_out, err.as<int32_t>(), st));
    else
      EDLGT_TRY(launch_mrows_cap<double>(cap, false, h, m, h.val.as<double>(), counts.as<int32_t>(), out,
                                         real_out, err.as<int32_t>(), st));
  }
  int32_t flag = 0;
  EDLGT_TRY(check_flag(err, st, &flag));
  if (flag) return set_error(EDLGT_ERR_INDEX_OVERFLOW, "momentum row exceeded its entry bound");
  EDLGT_TRY(exclusive_scan_i32_i64(counts.as<int32_t>(), out->row_ptr.as<int64_t>(), m.n_cols, st));
  int64_t nnz = 0;
  EDLGT_HIP(hipMemcpyAsync(&nnz, out->row_ptr.as<int64_t>() + m.n_cols, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  out->nnz = nnz;
  EDLGT_TRY(out->col.alloc((size_t)std::max<int64_t>(nnz, 1) * sizeof(int32_t)));
  EDLGT_TRY(out->val.alloc((size_t)std::max<int64_t>(nnz, 1) * value_bytes(out->value_type)));
  if (m.n_cols > 0) {
    if (cplx_h)
      EDLGT_TRY(launch_mrows_cap<double2>(cap, true, h, m, h.val.as<double2>(), nullptr, out, real_out,
                                          err.as<int32_t>(), st));
    else
      EDLGT_TRY(launch_mrows_cap<double>(cap, true, h, m, h.val.as<double>(), nullptr, out, real_out,
                                         err.as<int32_t>(), st));
  }
  EDLGT_TRY(check_flag(err, st, &flag));
  if (flag) return set_error(EDLGT_ERR_INDEX_OVERFLOW, "momentum row exceeded its entry bound");
  EDLGT_TRY(csr_prepare_stream(*out, st));
  return EDLGT_OK;
}

}  // namespace edlgt
