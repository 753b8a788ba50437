// Schmidt spectra of a sector state over a bipartition of the lattice, on the
// device (replaces the psi_matrix reshape and its SVD in
// QMB_state.entanglement_entropy / reduced_density_matrix,
// edlgt/modeling/qmb_state.py:160-291, with the partition maps of
// QuantumModel._get_partition, edlgt/models/quantum_model.py:715-791).
//
// In a symmetry sector the coefficient matrix psi[subsystem config,
// environment config] is block diagonal up to permutations: the connected
// components of the (subsystem, environment) pairs present in the sector are
// independent blocks (for a gauge theory, one per boundary flux and charge
// split).  The host lays the blocks out once per partition: block b is a
// dense d_b x c_b row-major matrix M_b at element offset moff_b (its rows the
// smaller side, or the subsystem side for a reduced density matrix), and
// basis state i goes to element pos[i].  On the device:
//   1. scatter psi into the packed blocks,
//   2. G_b = M_b M_b^H (16 x 16 LDS tiles, upper triangle computed and
//      mirrored, so G_b is exactly Hermitian),
//   3. eigenvalues of every G_b by cyclic two-sided Jacobi, all blocks at
//      once: each step applies d/2 disjoint rotations per block (round-robin
//      pairing), one workgroup per rotation, rows in one launch and columns in
//      the next; a sweep is d' - 1 steps (d' = d rounded up to even) and
//      sweeps repeat until every block's off-diagonal norm is below 1e-14 of
//      its diagonal (or stops falling below 1e-10).
// The squared singular values of M_b are the eigenvalues of G_b.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace edlgt {
namespace {

constexpr int kTileG = 16;
constexpr int kMaxSweeps = 60;

__device__ inline double2 c_of(double v) { return make_double2(v, 0.0); }
__device__ inline double2 c_of(double2 v) { return v; }

// M[pos[i]] = psi[i]; a position outside [0, m_elems) sets *bad instead
template <class S>
__global__ void schmidt_scatter(const S* __restrict__ psi, int64_t n, const int64_t* __restrict__ pos,
                                int64_t m_elems, double2* __restrict__ M, int* __restrict__ bad) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = pos[i];
    if (p < 0 || p >= m_elems) {
      *bad = 1;
      continue;
    }
    M[p] = c_of(psi[i]);
  }
}

// block table on the device: d, c, moff, goff, eoff (offset of its eigenvalues)
constexpr int kBlk = 5;

// One 16 x 16 tile (ti <= tj) of G_b = M_b M_b^H; the tile below the diagonal
// receives the conjugates.
__global__ void __launch_bounds__(256) schmidt_gram_tiles(const double2* __restrict__ M,
                                                          const int64_t* __restrict__ blk,
                                                          const int32_t* __restrict__ tiles,
                                                          double2* __restrict__ G) {
  __shared__ double2 A[kTileG][kTileG + 1], B[kTileG][kTileG + 1];
  const int b = tiles[3 * blockIdx.x], ti = tiles[3 * blockIdx.x + 1], tj = tiles[3 * blockIdx.x + 2];
  const int64_t d = blk[kBlk * b], c = blk[kBlk * b + 1], moff = blk[kBlk * b + 2], goff = blk[kBlk * b + 3];
  const int tx = threadIdx.x % kTileG, ty = threadIdx.x / kTileG;
  const int64_t i0 = (int64_t)ti * kTileG, j0 = (int64_t)tj * kTileG;
  double re = 0.0, im = 0.0;
  for (int64_t k0 = 0; k0 < c; k0 += kTileG) {
    const int64_t k = k0 + tx;
    A[ty][tx] = (i0 + ty < d && k < c) ? M[moff + (i0 + ty) * c + k] : make_double2(0.0, 0.0);
    B[ty][tx] = (j0 + ty < d && k < c) ? M[moff + (j0 + ty) * c + k] : make_double2(0.0, 0.0);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTileG; ++kk) {
      const double2 a = A[ty][kk], w = B[tx][kk];  // a * conj(w)
      re += a.x * w.x + a.y * w.y;
      im += a.y * w.x - a.x * w.y;
    }
    __syncthreads();
  }
  const int64_t i = i0 + ty, j = j0 + tx;
  if (i < d && j < d && i <= j) {
    if (i == j) {
      G[goff + i * d + i] = make_double2(re, 0.0);
    } else {
      G[goff + i * d + j] = make_double2(re, im);
      G[goff + j * d + i] = make_double2(re, -im);
    }
  }
}

// Rotation U = diag(1, conj(e)) [[c, s], [-s, c]], e = z / |z|, that zeroes
// G[p][q] = z of the Hermitian 2 x 2 block [[a, z], [conj z, b]] (Golub & Van
// Loan's symmetric Schur step after removing the phase of z).
struct Rot {
  double c, s, ex, ey, a, b, az, t;
  int32_t p, q, on, pad;
};

__device__ inline int rr_position(int j, int t, int np) { return j == 0 ? 0 : ((j - 1 + t) % (np - 1)) + 1; }

// Step t of a sweep, rows: for work item w = (block, k) the pair is the k-th
// of the round-robin schedule; its rotation is computed from the current G
// and rows p, q become U^H rows.
__global__ void __launch_bounds__(64) jacobi_rows(double2* __restrict__ G, const int64_t* __restrict__ blk,
                                                  const int32_t* __restrict__ work, int t, Rot* __restrict__ rot) {
  const int w = blockIdx.x;
  const int b = work[2 * w], k = work[2 * w + 1];
  const int64_t d = blk[kBlk * b], goff = blk[kBlk * b + 3];
  const int np = (int)(d + (d & 1));
  Rot r{};
  r.on = 0;
  double2* Gb = G + goff;
  if (t < np - 1) {
    int p = rr_position(k, t, np), q = rr_position(np - 1 - k, t, np);
    if (p > q) {
      const int x = p;
      p = q;
      q = x;
    }
    if (q < d) {
      const double a = Gb[(int64_t)p * d + p].x, bb = Gb[(int64_t)q * d + q].x;
      const double2 z = Gb[(int64_t)p * d + q];
      const double az = hypot(z.x, z.y);
      if (az > 0.0) {
        const double tau = (bb - a) / (2.0 * az);
        const double t1 = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        r.c = 1.0 / sqrt(1.0 + t1 * t1);
        r.s = t1 * r.c;
        r.ex = z.x / az;
        r.ey = z.y / az;
        r.a = a;
        r.b = bb;
        r.az = az;
        r.t = t1;
        r.p = p;
        r.q = q;
        r.on = 1;
      }
    }
  }
  __syncthreads();  // every lane has read the 2 x 2 block before rows change
  if (threadIdx.x == 0) rot[w] = r;
  if (!r.on) return;
  const double c = r.c, s = r.s, ex = r.ex, ey = r.ey;
  double2* rp = Gb + (int64_t)r.p * d;
  double2* rq = Gb + (int64_t)r.q * d;
  for (int64_t j = threadIdx.x; j < d; j += 64) {
    const double2 u = rp[j], v = rq[j];
    const double evx = ex * v.x - ey * v.y, evy = ex * v.y + ey * v.x;  // e * v
    rp[j] = make_double2(c * u.x - s * evx, c * u.y - s * evy);
    rq[j] = make_double2(s * u.x + c * evx, s * u.y + c * evy);
  }
}

// Step t, columns: columns p, q become (G U) columns; the 2 x 2 block is then
// set to its exact rotated values (real diagonal, zero coupling).
__global__ void __launch_bounds__(64) jacobi_cols(double2* __restrict__ G, const int64_t* __restrict__ blk,
                                                  const int32_t* __restrict__ work, const Rot* __restrict__ rot) {
  const Rot r = rot[blockIdx.x];
  if (!r.on) return;
  const int b = work[2 * blockIdx.x];
  const int64_t d = blk[kBlk * b], goff = blk[kBlk * b + 3];
  double2* Gb = G + goff;
  const double c = r.c, s = r.s, ex = r.ex, ey = -r.ey;  // conj(e)
  for (int64_t i = threadIdx.x; i < d; i += 64) {
    double2* row = Gb + i * d;
    const double2 u = row[r.p], v = row[r.q];
    const double evx = ex * v.x - ey * v.y, evy = ex * v.y + ey * v.x;  // conj(e) * v
    row[r.p] = make_double2(c * u.x - s * evx, c * u.y - s * evy);
    row[r.q] = make_double2(s * u.x + c * evx, s * u.y + c * evy);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Gb[(int64_t)r.p * d + r.p] = make_double2(r.a - r.t * r.az, 0.0);
    Gb[(int64_t)r.q * d + r.q] = make_double2(r.b + r.t * r.az, 0.0);
    Gb[(int64_t)r.p * d + r.q] = make_double2(0.0, 0.0);
    Gb[(int64_t)r.q * d + r.p] = make_double2(0.0, 0.0);
  }
}

// out[2b] = sum of |G_ij|^2 off the diagonal, out[2b+1] = sum of |G_ii|^2
__global__ void __launch_bounds__(256) block_norms(const double2* __restrict__ G, const int64_t* __restrict__ blk,
                                                   double* __restrict__ out) {
  __shared__ double so[256], sd[256];
  const int b = blockIdx.x;
  const int64_t d = blk[kBlk * b], goff = blk[kBlk * b + 3];
  double off = 0.0, dia = 0.0;
  for (int64_t e = threadIdx.x; e < d * d; e += 256) {
    const double2 g = G[goff + e];
    const double v = g.x * g.x + g.y * g.y;
    if (e / d == e % d) dia += v;
    else off += v;
  }
  so[threadIdx.x] = off;
  sd[threadIdx.x] = dia;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      so[threadIdx.x] += so[threadIdx.x + o];
      sd[threadIdx.x] += sd[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * b] = so[0];
    out[2 * b + 1] = sd[0];
  }
}

__global__ void block_diagonals(const double2* __restrict__ G, const int64_t* __restrict__ blk,
                                double* __restrict__ out) {
  const int b = blockIdx.x;
  const int64_t d = blk[kBlk * b], goff = blk[kBlk * b + 3], eoff = blk[kBlk * b + 4];
  for (int64_t i = threadIdx.x; i < d; i += blockDim.x) out[eoff + i] = G[goff + i * d + i].x;
}

// device table (d, c, moff, goff, eoff) from the caller's (d, c, moff, goff)
int block_table(int n_blocks, const int64_t* blocks, std::vector<int64_t>& tab, int64_t& m_elems, int64_t& g_elems,
                int64_t& n_vals) {
  tab.assign((size_t)kBlk * std::max(n_blocks, 1), 0);
  m_elems = g_elems = n_vals = 0;
  for (int b = 0; b < n_blocks; ++b) {
    const int64_t d = blocks[4 * b], c = blocks[4 * b + 1], mo = blocks[4 * b + 2], go = blocks[4 * b + 3];
    constexpr int64_t kLim = (int64_t)1 << 52;
    if (d < 1 || c < 1 || mo < 0 || go < 0 || d > ((int64_t)1 << 26) || c > kLim / d || mo > kLim - d * c ||
        go > kLim - d * d)
      return set_error(EDLGT_ERR_INVALID_ARG, "block " + std::to_string(b) + ": bad shape or offset");
    m_elems = std::max(m_elems, mo + d * c);
    g_elems = std::max(g_elems, go + d * d);
    tab[kBlk * b] = d;
    tab[kBlk * b + 1] = c;
    tab[kBlk * b + 2] = mo;
    tab[kBlk * b + 3] = go;
    tab[kBlk * b + 4] = n_vals;
    n_vals += d;
  }
  return EDLGT_OK;
}

}  // namespace

int schmidt_gram(const void* psi, bool cplx, int64_t n, const int64_t* pos, int n_blocks, const int64_t* blocks,
                 void* gram, int64_t gram_elems, hipStream_t st) {
  std::vector<int64_t> tab;
  int64_t m_elems = 0, g_elems = 0, n_vals = 0;
  EDLGT_TRY(block_table(n_blocks, blocks, tab, m_elems, g_elems, n_vals));
  if (g_elems > gram_elems)
    return set_error(EDLGT_ERR_INVALID_ARG, "the Gram blocks need " + std::to_string(g_elems) + " elements, the buffer holds " +
                                                std::to_string(gram_elems));
  if (n_blocks == 0) return EDLGT_OK;
  std::vector<int32_t> tiles;
  for (int b = 0; b < n_blocks; ++b) {
    const int nt = (int)((tab[kBlk * b] + kTileG - 1) / kTileG);
    for (int i = 0; i < nt; ++i)
      for (int j = i; j < nt; ++j) {
        tiles.push_back(b);
        tiles.push_back(i);
        tiles.push_back(j);
      }
  }
  DevBuf dtab, dtiles, M, bad;
  EDLGT_TRY(upload(dtab, tab.data(), tab.size(), st));
  EDLGT_TRY(upload(dtiles, tiles.data(), tiles.size(), st));
  EDLGT_TRY(M.alloc((size_t)m_elems * sizeof(double2)));
  EDLGT_TRY(bad.alloc(sizeof(int)));
  EDLGT_HIP(hipMemsetAsync(M.p, 0, (size_t)m_elems * sizeof(double2), st));
  EDLGT_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), st));
  if (n > 0) {
    const unsigned g = (unsigned)std::min<int64_t>(65536, (n + 255) / 256);
    if (cplx)
      hipLaunchKernelGGL(schmidt_scatter<double2>, dim3(g), dim3(256), 0, st, (const double2*)psi, n, pos, m_elems,
                         M.as<double2>(), bad.as<int>());
    else
      hipLaunchKernelGGL(schmidt_scatter<double>, dim3(g), dim3(256), 0, st, (const double*)psi, n, pos, m_elems,
                         M.as<double2>(), bad.as<int>());
    EDLGT_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(schmidt_gram_tiles, dim3((unsigned)(tiles.size() / 3)), dim3(256), 0, st, M.as<double2>(),
                     dtab.as<int64_t>(), dtiles.as<int32_t>(), (double2*)gram);
  EDLGT_HIP(hipGetLastError());
  int hbad = 0;
  EDLGT_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  if (hbad) return set_error(EDLGT_ERR_INVALID_ARG, "a state position lies outside the blocks");
  return EDLGT_OK;
}

int hermitian_eigvals(void* gram, int64_t gram_elems, int n_blocks, const int64_t* blocks, double* evals,
                      int64_t n_evals, hipStream_t st) {
  std::vector<int64_t> tab;
  int64_t m_elems = 0, g_elems = 0, n_vals = 0;
  EDLGT_TRY(block_table(n_blocks, blocks, tab, m_elems, g_elems, n_vals));
  if (g_elems > gram_elems || n_vals > n_evals)
    return set_error(EDLGT_ERR_INVALID_ARG, "Gram buffer or eigenvalue array smaller than the blocks");
  if (n_blocks == 0) return EDLGT_OK;
  std::vector<int32_t> work;
  int max_np = 0;
  for (int b = 0; b < n_blocks; ++b) {
    const int d = (int)tab[kBlk * b];
    if (d < 2) continue;
    const int np = d + (d & 1);
    max_np = std::max(max_np, np);
    for (int k = 0; k < np / 2; ++k) {
      work.push_back(b);
      work.push_back(k);
    }
  }
  DevBuf dtab, dwork, rot, norms, vals;
  EDLGT_TRY(upload(dtab, tab.data(), tab.size(), st));
  EDLGT_TRY(norms.alloc((size_t)2 * n_blocks * sizeof(double)));
  EDLGT_TRY(vals.alloc((size_t)n_vals * sizeof(double)));
  double2* G = (double2*)gram;
  const int nw = (int)(work.size() / 2);
  if (nw > 0) {
    EDLGT_TRY(upload(dwork, work.data(), work.size(), st));
    EDLGT_TRY(rot.alloc((size_t)nw * sizeof(Rot)));
    std::vector<double> hn((size_t)2 * n_blocks);
    std::vector<double> prev((size_t)n_blocks, INFINITY);
    bool done = false;
    for (int sweep = 0; sweep < kMaxSweeps && !done; ++sweep) {
      for (int t = 0; t < max_np - 1; ++t) {
        hipLaunchKernelGGL(jacobi_rows, dim3(nw), dim3(64), 0, st, G, dtab.as<int64_t>(), dwork.as<int32_t>(), t,
                           rot.as<Rot>());
        hipLaunchKernelGGL(jacobi_cols, dim3(nw), dim3(64), 0, st, G, dtab.as<int64_t>(), dwork.as<int32_t>(),
                           rot.as<const Rot>());
      }
      EDLGT_HIP(hipGetLastError());
      hipLaunchKernelGGL(block_norms, dim3(n_blocks), dim3(256), 0, st, G, dtab.as<int64_t>(), norms.as<double>());
      EDLGT_HIP(hipGetLastError());
      EDLGT_HIP(hipMemcpyAsync(hn.data(), norms.p, hn.size() * sizeof(double), hipMemcpyDeviceToHost, st));
      EDLGT_HIP(hipStreamSynchronize(st));
      done = true;
      for (int b = 0; b < n_blocks; ++b) {
        const double rel = std::sqrt(hn[2 * b] / std::max(hn[2 * b + 1], 1e-300));
        const bool stalled = rel <= 1e-10 && !(rel < 0.5 * prev[b]);
        if (!(rel <= 1e-14 || stalled || hn[2 * b] == 0.0)) done = false;
        prev[b] = rel;
      }
    }
    if (!done) return set_error(EDLGT_ERR_NOT_CONVERGED, "Jacobi sweeps did not converge");
  }
  hipLaunchKernelGGL(block_diagonals, dim3(n_blocks), dim3(256), 0, st, G, dtab.as<int64_t>(), vals.as<double>());
  EDLGT_HIP(hipGetLastError());
  EDLGT_HIP(hipMemcpyAsync(evals, vals.p, (size_t)n_vals * sizeof(double), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  return EDLGT_OK;
}

}  // namespace edlgt
