// Symmetry-sector enumeration on the device.
//
// The sector is the set of configurations c (c_s < loc_dims[s]) that satisfy
// the global / link / n-body constraints, in lexicographic order with site 0
// most significant (the row order of the reference's sector_configs,
// edlgt/symmetries/symmetry_sector.py:785-967 and :1329-1469).
//
// Design (MI355X):
//  1. Breadth-first frontier of valid prefixes.  A prefix of depth p is its
//     mixed-radix index over sites [0, p).  Expanding a frontier by one site
//     launches one thread per candidate (prefix*d_p + s), checks only the
//     constraints that become decidable at site p (U(1) reachability with
//     suffix min/max bounds, link pairs and n-body groups whose last site is
//     p), and compacts the survivors with a flag scan.  The expansion keeps
//     prefix order, so the frontier stays sorted.
//  2. Once the frontier holds enough prefixes to fill the chip (or the next
//     candidate set would be too large), every frontier prefix is completed
//     by one thread with an iterative depth-first search over the remaining
//     sites, in lexicographic order: a count pass, an int64 scan of the
//     counts, and a fill pass that writes keys and configuration bytes at
//     the scanned offsets.  Output order = (prefix order, DFS order) =
//     lexicographic order, with no global sort.
// Integer output is deterministic and independent of launch shapes.
#include <algorithm>
#include <climits>
#include <cmath>

#include "common.h"

namespace edlgt {
namespace {

constexpr int kMaxSites = 48;
constexpr int kMaxGlob = 4;
constexpr double kSymTol = 1e-13;  // symmetry_sector.py:17

struct SectorDev {
  int n_sites, d_max;
  int n_glob, glob_type;
  const double* glob_diags;    // [n_glob][n_sites][d_max]
  const double* glob_sectors;  // [n_glob]
  const double* rem_min;       // [n_glob][n_sites+1]  (U1 suffix bounds)
  const double* rem_max;
  const int32_t* act_link_off; // [n_sites+1]
  const int32_t* act_link;     // triples (dir, i, j) activated at their max site
  const double* link_diags;    // [n_dirs][2][n_sites][d_max]
  const double* link_sectors;  // [n_dirs]
  const int32_t* act_nb_off;   // [n_sites+1]
  const int32_t* act_nb;       // n-body ids activated at their max site
  int nbody_type;
  const double* nbody_diags;   // [n_nb][n_sites][d_max]
  const double* nbody_sectors;
  const int32_t* nb_site_off;  // [n_nb+1]
  const int32_t* nb_sites;
  const int32_t* loc_dims;     // [n_sites]
  const uint64_t* strides;     // [n_sites]
  // link_mask[k * d_max + v] (d_max <= 64, else null): bit a set iff the
  // activated link k holds with its later site in state a and its other site
  // in state v
  const uint64_t* link_mask;
};

// n-body constraint `id` on a full assignment of its sites: additive (U1),
// multiplicative (Z2) or additive modulo m (type EDLGT_SYM_MOD(m), integer values).
__device__ inline bool nbody_ok(const SectorDev& D, int id, const uint8_t* dig) {
  const int n = D.n_sites;
  const bool mult = D.nbody_type == EDLGT_SYM_Z2;
  double v = mult ? 1.0 : 0.0;
  for (int q = D.nb_site_off[id]; q < D.nb_site_off[id + 1]; ++q) {
    const int site = D.nb_sites[q];
    const double x = D.nbody_diags[((size_t)id * n + site) * D.d_max + dig[site]];
    v = mult ? v * x : v + x;
  }
  if (D.nbody_type >= 16) {
    const long long m = D.nbody_type - 16;
    long long r = (llround(v) - llround(D.nbody_sectors[id])) % m;
    return r == 0;
  }
  return fabs(v - D.nbody_sectors[id]) <= kSymTol;
}

// Checks every constraint that becomes decidable when site s is assigned.
// gin: global partial values of sites [0, s); gout receives those of [0, s].
__device__ inline bool check_site(const SectorDev& D, int s, const uint8_t* dig,
                                  const double* gin, double* gout) {
  const int n = D.n_sites;
  const int a = dig[s];
  for (int g = 0; g < D.n_glob; ++g) {
    const double v = D.glob_diags[((size_t)g * n + s) * D.d_max + a];
    if (D.glob_type == EDLGT_SYM_U1) {
      const double nv = (s == 0 ? 0.0 : gin[g]) + v;
      gout[g] = nv;
      const double t = D.glob_sectors[g];
      if (s < n - 1) {
        const double lo = nv + D.rem_min[g * (n + 1) + s + 1];
        const double hi = nv + D.rem_max[g * (n + 1) + s + 1];
        if (t < lo - kSymTol || t > hi + kSymTol) return false;
      } else if (fabs(nv - t) > kSymTol) {
        return false;
      }
    } else {
      const double nv = (s == 0 ? 1.0 : gin[g]) * v;
      gout[g] = nv;
      if (s == n - 1 && fabs(nv - D.glob_sectors[g]) > kSymTol) return false;
    }
  }
  for (int k = D.act_link_off[s]; k < D.act_link_off[s + 1]; ++k) {
    const int dir = D.act_link[3 * k], i = D.act_link[3 * k + 1], j = D.act_link[3 * k + 2];
    const double v = D.link_diags[(((size_t)dir * 2 + 0) * n + i) * D.d_max + dig[i]] +
                     D.link_diags[(((size_t)dir * 2 + 1) * n + j) * D.d_max + dig[j]];
    if (fabs(v - D.link_sectors[dir]) > kSymTol) return false;
  }
  for (int k = D.act_nb_off[s]; k < D.act_nb_off[s + 1]; ++k)
    if (!nbody_ok(D, D.act_nb[k], dig)) return false;
  return true;
}

// Decodes a depth-p prefix index into digits [0, p).
__device__ inline void decode_prefix(const SectorDev& D, uint64_t idx, int p, uint8_t* dig) {
  for (int s = p - 1; s >= 0; --s) {
    const uint64_t d = (uint64_t)D.loc_dims[s];
    dig[s] = (uint8_t)(idx % d);
    idx /= d;
  }
}

// Re-accumulates the global partial values of a (valid) prefix of depth p.
__device__ inline void prefix_globals(const SectorDev& D, const uint8_t* dig, int p, double* g) {
  for (int k = 0; k < D.n_glob; ++k) {
    double v = D.glob_type == EDLGT_SYM_U1 ? 0.0 : 1.0;
    for (int s = 0; s < p; ++s) {
      const double x = D.glob_diags[((size_t)k * D.n_sites + s) * D.d_max + dig[s]];
      v = D.glob_type == EDLGT_SYM_U1 ? v + x : v * x;
    }
    g[k] = v;
  }
}

// One thread per candidate prefix*d_p + s of the next frontier level.
__global__ void expand_flags(SectorDev D, const uint64_t* __restrict__ frontier, int64_t n_front,
                             int p, int32_t* __restrict__ flags) {
  const int64_t dp = D.loc_dims[p];
  const int64_t n_cand = n_front * dp;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_cand;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = t / dp;
    const int s = (int)(t - f * dp);
    uint8_t dig[kMaxSites];
    decode_prefix(D, frontier[f], p, dig);
    dig[p] = (uint8_t)s;
    double gin[kMaxGlob], gout[kMaxGlob];
    prefix_globals(D, dig, p, gin);
    flags[t] = check_site(D, p, dig, gin, gout) ? 1 : 0;
  }
}

__global__ void expand_compact(const uint64_t* __restrict__ frontier, int64_t n_front, int64_t dp,
                               const int32_t* __restrict__ flags, const int64_t* __restrict__ pos,
                               uint64_t* __restrict__ next) {
  const int64_t n_cand = n_front * dp;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_cand;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (flags[t]) {
      const int64_t f = t / dp;
      next[pos[t]] = frontier[f] * (uint64_t)dp + (uint64_t)(t - f * dp);
    }
  }
}

// Depth-first completion of one frontier prefix.  kFill=false counts the
// completions, kFill=true writes them (keys + configuration bytes) from
// offsets[f] on, in lexicographic order.
template <bool kFill>
__global__ void complete_prefixes(SectorDev D, const uint64_t* __restrict__ frontier,
                                  int64_t n_front, int p, int64_t* __restrict__ counts,
                                  const int64_t* __restrict__ offsets, uint64_t* __restrict__ keys,
                                  uint8_t* __restrict__ configs, int cfg_stride) {
  const int n = D.n_sites;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < n_front;
       f += (int64_t)gridDim.x * blockDim.x) {
    uint8_t dig[kMaxSites];
    double gsum[kMaxSites + 1][kMaxGlob];  // gsum[s] = globals of sites [0, s)
    decode_prefix(D, frontier[f], p, dig);
    prefix_globals(D, dig, p, gsum[p]);
    uint64_t pkey = 0;
    for (int s = 0; s < p; ++s) pkey += (uint64_t)dig[s] * D.strides[s];
    int64_t count = 0;
    int64_t out = kFill ? offsets[f] : 0;
    if (p == n) {
      // The prefix already is a full configuration (validated by expansion).
      count = 1;
      if (kFill) {
        keys[out] = pkey;
        for (int s = 0; s < n; ++s) configs[out * cfg_stride + s] = dig[s];
      }
    } else {
      uint64_t kpart[kMaxSites + 1];
      kpart[p] = pkey;
      int s = p;
      int cur = -1;
      while (true) {
        ++cur;
        if (cur >= D.loc_dims[s]) {
          if (s == p) break;
          --s;
          cur = dig[s];
          continue;
        }
        dig[s] = (uint8_t)cur;
        if (!check_site(D, s, dig, gsum[s], gsum[s + 1])) continue;
        const uint64_t k = kpart[s] + (uint64_t)cur * D.strides[s];
        if (s == n - 1) {
          if (kFill) {
            keys[out] = k;
            uint8_t* row = configs + out * cfg_stride;
            for (int q = 0; q < n; ++q) row[q] = dig[q];
            ++out;
          }
          ++count;
        } else {
          kpart[s + 1] = k;
          ++s;
          cur = -1;
        }
      }
    }
    if (!kFill) counts[f] = count;
  }
}

// Depth-first completion with all per-level state outside private memory:
// digits live in LDS (64 bytes x 64 lanes), the global U(1) partial sums are
// exact integers (diagonal values scaled by 2^k on the host) kept in registers
// and updated incrementally on descent / backtrack, and the mixed-radix key is
// updated the same way.  Used whenever the global constraints are U(1) with
// dyadic-rational diagonal values (every model of this engine); the generic
// kernel above covers the rest.
struct IntGlob {
  const int64_t* tab;     // [n_glob][n_sites][d_max]
  const int64_t* rmin;    // [n_glob][n_sites+1]
  const int64_t* rmax;
  const int64_t* target;  // [n_glob]
};

__device__ inline bool check_links_nb(const SectorDev& D, int s, const uint8_t* dig) {
  const int n = D.n_sites;
  for (int k = D.act_link_off[s]; k < D.act_link_off[s + 1]; ++k) {
    const int dir = D.act_link[3 * k], i = D.act_link[3 * k + 1], j = D.act_link[3 * k + 2];
    const double v = D.link_diags[(((size_t)dir * 2 + 0) * n + i) * D.d_max + dig[i]] +
                     D.link_diags[(((size_t)dir * 2 + 1) * n + j) * D.d_max + dig[j]];
    if (fabs(v - D.link_sectors[dir]) > kSymTol) return false;
  }
  for (int k = D.act_nb_off[s]; k < D.act_nb_off[s + 1]; ++k)
    if (!nbody_ok(D, D.act_nb[k], dig)) return false;
  return true;
}

// States of site s that satisfy every link activated at s, given the
// digits of the earlier sites (one table lookup per link).
__device__ inline uint64_t allowed_states(const SectorDev& D, int s, const uint8_t* dig) {
  const int ds = D.loc_dims[s];
  uint64_t m = ds >= 64 ? ~0ull : ((1ull << ds) - 1);
  for (int k = D.act_link_off[s]; k < D.act_link_off[s + 1]; ++k) {
    const int i = D.act_link[3 * k + 1], j = D.act_link[3 * k + 2];
    const int other = i == s ? j : i;  // i == j == s: the table row does not depend on v
    m &= D.link_mask[(size_t)k * D.d_max + dig[other]];
  }
  return m;
}

__device__ inline bool check_nb(const SectorDev& D, int s, const uint8_t* dig) {
  for (int k = D.act_nb_off[s]; k < D.act_nb_off[s + 1]; ++k)
    if (!nbody_ok(D, D.act_nb[k], dig)) return false;
  return true;
}

template <bool kFill, class KeyT>
__global__ void __launch_bounds__(64) complete_prefixes_int(
    SectorDev D, IntGlob G, const KeyT* __restrict__ strides, const uint64_t* __restrict__ frontier,
    int64_t n_front, int p, int64_t* __restrict__ counts, const int64_t* __restrict__ offsets,
    KeyT* __restrict__ keys, uint8_t* __restrict__ configs, int cfg_stride) {
  __shared__ uint8_t sdig[64 * 64];
  uint8_t* dig = sdig + threadIdx.x * 64;
  const int n = D.n_sites;
  const int ng = D.n_glob;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < n_front;
       f += (int64_t)gridDim.x * blockDim.x) {
    decode_prefix(D, frontier[f], p, dig);
    int64_t g[kMaxGlob];
    KeyT key = 0;
#pragma unroll
    for (int k = 0; k < kMaxGlob; ++k) g[k] = 0;
    for (int s = 0; s < p; ++s) {
      key += (KeyT)dig[s] * strides[s];
#pragma unroll
      for (int k = 0; k < kMaxGlob; ++k)
        if (k < ng) g[k] += G.tab[((size_t)k * n + s) * D.d_max + dig[s]];
    }
    int64_t count = 0;
    int64_t out = kFill ? offsets[f] : 0;
    if (p == n) {
      count = 1;
      if (kFill) {
        keys[out] = key;
        for (int q = 0; q < n; ++q) configs[out * cfg_stride + q] = dig[q];
      }
    } else {
      int s = p;
      int cur = -1;
      const bool masks = D.link_mask != nullptr;
      while (true) {
        if (masks) {
          // next state above cur that the links activated at s allow
          const uint64_t m = allowed_states(D, s, dig) & (cur >= 63 ? 0ull : (~0ull << (cur + 1)));
          cur = m ? __ffsll((long long)m) - 1 : D.loc_dims[s];
        } else {
          ++cur;
        }
        if (cur >= D.loc_dims[s]) {
          if (s == p) break;
          --s;
          cur = dig[s];
          key -= (KeyT)cur * strides[s];
#pragma unroll
          for (int k = 0; k < kMaxGlob; ++k)
            if (k < ng) g[k] -= G.tab[((size_t)k * n + s) * D.d_max + cur];
          continue;
        }
        dig[s] = (uint8_t)cur;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < kMaxGlob; ++k) {
          if (k < ng && ok) {
            const int64_t gs = g[k] + G.tab[((size_t)k * n + s) * D.d_max + cur];
            const int64_t t = G.target[k];
            if (s < n - 1)
              ok = t >= gs + G.rmin[k * (n + 1) + s + 1] && t <= gs + G.rmax[k * (n + 1) + s + 1];
            else
              ok = gs == t;
          }
        }
        if (!ok || !(masks ? check_nb(D, s, dig) : check_links_nb(D, s, dig))) continue;
        if (s == n - 1) {
          if (kFill) {
            keys[out] = key + (KeyT)cur * strides[s];
            uint8_t* row = configs + out * cfg_stride;
            for (int q = 0; q < n; ++q) row[q] = dig[q];
            ++out;
          }
          ++count;
        } else {
          key += (KeyT)cur * strides[s];
#pragma unroll
          for (int k = 0; k < kMaxGlob; ++k)
            if (k < ng) g[k] += G.tab[((size_t)k * n + s) * D.d_max + cur];
          ++s;
          cur = -1;
        }
      }
    }
    if (!kFill) counts[f] = count;
  }
}

}  // namespace

int sector_build(int32_t n_sites, const int32_t* loc_dims, int32_t d_max, int32_t n_glob,
                 int32_t glob_type, const double* glob_diags, const double* glob_sectors,
                 int32_t n_dirs, const double* link_diags, const double* link_sectors,
                 int32_t n_pairs, const int32_t* pair_dir, const int32_t* pair_site0,
                 const int32_t* pair_site1, int32_t n_nbody, int32_t nbody_type,
                 const double* nbody_diags, const double* nbody_sectors,
                 const int32_t* nb_site_off_h, const int32_t* nb_sites_h, hipStream_t st,
                 Basis* out) {
  if (n_sites < 1 || n_sites > kMaxSites)
    return set_error(EDLGT_ERR_INVALID_ARG, "n_sites must be in [1, 48]");
  if (d_max < 1 || n_glob < 0 || n_glob > kMaxGlob || n_pairs < 0 || n_nbody < 0 || n_dirs < 0)
    return set_error(EDLGT_ERR_INVALID_ARG, "invalid constraint counts");
  if (n_glob && (glob_type != EDLGT_SYM_U1 && glob_type != EDLGT_SYM_Z2))
    return set_error(EDLGT_ERR_INVALID_ARG, "glob_type must be U1 or Z2");
  // Mixed-radix strides (last site fastest); keys are 64-bit when the whole
  // product space fits, 128-bit otherwise.
  std::vector<uint64_t> strides(n_sites);
  std::vector<u128> strides128(n_sites);
  u128 prodd = 1;
  bool wide = false;
  for (int s = n_sites - 1; s >= 0; --s) {
    if (loc_dims[s] < 1 || loc_dims[s] > 255 || loc_dims[s] > d_max)
      return set_error(EDLGT_ERR_INVALID_ARG, "loc_dims must be in [1, min(255, d_max)]");
    strides128[s] = prodd;
    strides[s] = (uint64_t)prodd;
    if (prodd > (((u128)1) << 120) / 256)
      return set_error(EDLGT_ERR_INDEX_OVERFLOW, "prod(loc_dims) does not fit 128-bit keys");
    prodd *= (unsigned)loc_dims[s];
    if (prodd > (u128)UINT64_MAX) wide = true;
  }
  // Host-side tables: U(1) suffix bounds, activation lists.
  std::vector<double> rem_min((size_t)std::max(n_glob, 1) * (n_sites + 1), 0.0),
      rem_max(rem_min.size(), 0.0);
  if (glob_type == EDLGT_SYM_U1) {
    for (int g = 0; g < n_glob; ++g)
      for (int s = n_sites - 1; s >= 0; --s) {
        const double* v = glob_diags + ((size_t)g * n_sites + s) * d_max;
        double mn = v[0], mx = v[0];
        for (int a = 1; a < loc_dims[s]; ++a) { mn = std::min(mn, v[a]); mx = std::max(mx, v[a]); }
        rem_min[g * (n_sites + 1) + s] = rem_min[g * (n_sites + 1) + s + 1] + mn;
        rem_max[g * (n_sites + 1) + s] = rem_max[g * (n_sites + 1) + s + 1] + mx;
      }
  }
  std::vector<std::vector<int32_t>> by_site_link(n_sites), by_site_nb(n_sites);
  for (int q = 0; q < n_pairs; ++q) {
    const int d = pair_dir[q], i = pair_site0[q], j = pair_site1[q];
    if (d < 0 || d >= n_dirs || i < 0 || j < 0 || i >= n_sites || j >= n_sites)
      return set_error(EDLGT_ERR_INVALID_ARG, "link pair out of range");
    auto& v = by_site_link[std::max(i, j)];
    v.push_back(d); v.push_back(i); v.push_back(j);
  }
  if (n_nbody > 0 && nbody_type >= 16) {
    if (nbody_type - 16 < 2) return set_error(EDLGT_ERR_INVALID_ARG, "modulus must be >= 2");
    for (size_t i = 0; i < (size_t)n_nbody * n_sites * d_max; ++i)
      if (std::fabs(nbody_diags[i] - std::nearbyint(nbody_diags[i])) > 1e-9)
        return set_error(EDLGT_ERR_INVALID_ARG, "modular n-body constraints need integer values");
  }
  for (int k = 0; k < n_nbody; ++k) {
    int mx = -1;
    for (int q = nb_site_off_h[k]; q < nb_site_off_h[k + 1]; ++q) {
      if (nb_sites_h[q] < 0 || nb_sites_h[q] >= n_sites)
        return set_error(EDLGT_ERR_INVALID_ARG, "n-body site out of range");
      mx = std::max(mx, nb_sites_h[q]);
    }
    if (mx >= 0) by_site_nb[mx].push_back(k);
  }
  std::vector<int32_t> act_link_off(n_sites + 1, 0), act_link, act_nb_off(n_sites + 1, 0), act_nb;
  for (int s = 0; s < n_sites; ++s) {
    act_link.insert(act_link.end(), by_site_link[s].begin(), by_site_link[s].end());
    act_link_off[s + 1] = (int32_t)(act_link.size() / 3);
    act_nb.insert(act_nb.end(), by_site_nb[s].begin(), by_site_nb[s].end());
    act_nb_off[s + 1] = (int32_t)act_nb.size();
  }
  // Link masks for the depth-first completion (states of a link's later
  // site allowed by each state of its other site), same test as check_site.
  std::vector<uint64_t> link_mask;
  const int n_act = (int)(act_link.size() / 3);
  if (d_max <= 64 && n_act > 0) {
    link_mask.assign((size_t)n_act * d_max, 0);
    for (int s = 0; s < n_sites; ++s)
      for (int k = act_link_off[s]; k < act_link_off[s + 1]; ++k) {
        const int dir = act_link[3 * k], i = act_link[3 * k + 1], j = act_link[3 * k + 2];
        for (int v = 0; v < d_max; ++v) {
          uint64_t m = 0;
          for (int a = 0; a < loc_dims[s]; ++a) {
            // the later site s takes state a, the other site state v (i == j: both a)
            const int ai = i == s ? a : v, aj = j == s ? a : v;
            const double x = link_diags[(((size_t)dir * 2 + 0) * n_sites + i) * d_max + ai] +
                             link_diags[(((size_t)dir * 2 + 1) * n_sites + j) * d_max + aj];
            if (std::fabs(x - link_sectors[dir]) <= kSymTol) m |= 1ull << a;
          }
          link_mask[(size_t)k * d_max + v] = m;
        }
      }
  }
  // Upload.
  DevBuf b_gd, b_gs, b_rmin, b_rmax, b_alo, b_al, b_ld, b_ls, b_ano, b_an, b_nd, b_ns, b_nso, b_nsi,
      b_ldims, b_str, b_lmask;
  if (!link_mask.empty()) EDLGT_TRY(upload(b_lmask, link_mask.data(), link_mask.size(), st));
  const double zero = 0.0;
  const int32_t izero = 0;
  EDLGT_TRY(upload(b_gd, n_glob ? glob_diags : &zero, n_glob ? (size_t)n_glob * n_sites * d_max : 1, st));
  EDLGT_TRY(upload(b_gs, n_glob ? glob_sectors : &zero, n_glob ? (size_t)n_glob : 1, st));
  EDLGT_TRY(upload(b_rmin, rem_min.data(), rem_min.size(), st));
  EDLGT_TRY(upload(b_rmax, rem_max.data(), rem_max.size(), st));
  EDLGT_TRY(upload(b_alo, act_link_off.data(), act_link_off.size(), st));
  EDLGT_TRY(upload(b_al, act_link.empty() ? &izero : act_link.data(), std::max<size_t>(act_link.size(), 1), st));
  EDLGT_TRY(upload(b_ld, n_dirs ? link_diags : &zero, n_dirs ? (size_t)n_dirs * 2 * n_sites * d_max : 1, st));
  EDLGT_TRY(upload(b_ls, n_dirs ? link_sectors : &zero, n_dirs ? (size_t)n_dirs : 1, st));
  EDLGT_TRY(upload(b_ano, act_nb_off.data(), act_nb_off.size(), st));
  EDLGT_TRY(upload(b_an, act_nb.empty() ? &izero : act_nb.data(), std::max<size_t>(act_nb.size(), 1), st));
  EDLGT_TRY(upload(b_nd, n_nbody ? nbody_diags : &zero, n_nbody ? (size_t)n_nbody * n_sites * d_max : 1, st));
  EDLGT_TRY(upload(b_ns, n_nbody ? nbody_sectors : &zero, n_nbody ? (size_t)n_nbody : 1, st));
  EDLGT_TRY(upload(b_nso, n_nbody ? nb_site_off_h : &izero, n_nbody ? (size_t)n_nbody + 1 : 1, st));
  const size_t n_nbs = n_nbody ? (size_t)nb_site_off_h[n_nbody] : 0;
  EDLGT_TRY(upload(b_nsi, n_nbs ? nb_sites_h : &izero, std::max<size_t>(n_nbs, 1), st));
  EDLGT_TRY(upload(b_ldims, loc_dims, n_sites, st));
  EDLGT_TRY(upload(b_str, strides.data(), strides.size(), st));

  SectorDev D;
  D.n_sites = n_sites; D.d_max = d_max; D.n_glob = n_glob; D.glob_type = glob_type;
  D.glob_diags = b_gd.as<double>(); D.glob_sectors = b_gs.as<double>();
  D.rem_min = b_rmin.as<double>(); D.rem_max = b_rmax.as<double>();
  D.act_link_off = b_alo.as<int32_t>(); D.act_link = b_al.as<int32_t>();
  D.link_diags = b_ld.as<double>(); D.link_sectors = b_ls.as<double>();
  D.act_nb_off = b_ano.as<int32_t>(); D.act_nb = b_an.as<int32_t>();
  D.nbody_type = nbody_type; D.nbody_diags = b_nd.as<double>(); D.nbody_sectors = b_ns.as<double>();
  D.nb_site_off = b_nso.as<int32_t>(); D.nb_sites = b_nsi.as<int32_t>();
  D.loc_dims = b_ldims.as<int32_t>(); D.strides = b_str.as<uint64_t>();
  D.link_mask = link_mask.empty() ? nullptr : b_lmask.as<uint64_t>();

  // Exact integer form of the U(1) constraints: every diagonal value and
  // target scaled by 2^k (k <= 12) must be an integer of modest size.
  bool int_mode = (n_glob == 0 || glob_type == EDLGT_SYM_U1) && n_sites <= 64;
  std::vector<int64_t> gtab_i, rmin_i, rmax_i, tgt_i;
  if (int_mode && n_glob > 0) {
    int scale_pow = -1;
    for (int k = 0; k <= 12 && scale_pow < 0; ++k) {
      const double sc = std::ldexp(1.0, k);
      bool ok = true;
      for (int g = 0; g < n_glob && ok; ++g) {
        ok = std::fabs(glob_sectors[g] * sc - std::nearbyint(glob_sectors[g] * sc)) < 1e-9;
        for (int s = 0; s < n_sites && ok; ++s)
          for (int a = 0; a < loc_dims[s] && ok; ++a) {
            const double v = glob_diags[((size_t)g * n_sites + s) * d_max + a] * sc;
            ok = std::fabs(v - std::nearbyint(v)) < 1e-9 && std::fabs(v) < 1e12;
          }
      }
      if (ok) scale_pow = k;
    }
    int_mode = scale_pow >= 0;
    if (int_mode) {
      const double sc = std::ldexp(1.0, scale_pow);
      gtab_i.assign((size_t)n_glob * n_sites * d_max, 0);
      rmin_i.assign((size_t)n_glob * (n_sites + 1), 0);
      rmax_i.assign((size_t)n_glob * (n_sites + 1), 0);
      tgt_i.assign(n_glob, 0);
      for (int g = 0; g < n_glob; ++g) {
        tgt_i[g] = (int64_t)std::llround(glob_sectors[g] * sc);
        for (int s = n_sites - 1; s >= 0; --s) {
          int64_t mn = INT64_MAX, mx = INT64_MIN;
          for (int a = 0; a < loc_dims[s]; ++a) {
            const int64_t v = (int64_t)std::llround(glob_diags[((size_t)g * n_sites + s) * d_max + a] * sc);
            gtab_i[((size_t)g * n_sites + s) * d_max + a] = v;
            mn = std::min(mn, v);
            mx = std::max(mx, v);
          }
          rmin_i[g * (n_sites + 1) + s] = rmin_i[g * (n_sites + 1) + s + 1] + mn;
          rmax_i[g * (n_sites + 1) + s] = rmax_i[g * (n_sites + 1) + s + 1] + mx;
        }
      }
    }
  }
  if (wide && !int_mode)
    return set_error(EDLGT_ERR_INVALID_ARG,
                     "128-bit keys need additive global constraints with dyadic values");
  DevBuf b_gti, b_rmi, b_rxi, b_tgi;
  IntGlob IG{nullptr, nullptr, nullptr, nullptr};
  if (int_mode) {
    const int64_t z = 0;
    EDLGT_TRY(upload(b_gti, gtab_i.empty() ? &z : gtab_i.data(), std::max<size_t>(gtab_i.size(), 1), st));
    EDLGT_TRY(upload(b_rmi, rmin_i.empty() ? &z : rmin_i.data(), std::max<size_t>(rmin_i.size(), 1), st));
    EDLGT_TRY(upload(b_rxi, rmax_i.empty() ? &z : rmax_i.data(), std::max<size_t>(rmax_i.size(), 1), st));
    EDLGT_TRY(upload(b_tgi, tgt_i.empty() ? &z : tgt_i.data(), std::max<size_t>(tgt_i.size(), 1), st));
    IG = IntGlob{b_gti.as<int64_t>(), b_rmi.as<int64_t>(), b_rxi.as<int64_t>(), b_tgi.as<int64_t>()};
  }

  // Phase 1: breadth-first frontier.
  const int64_t kTargetFrontier = 1 << 20;
  const int64_t kMaxCandidates = (int64_t)1 << 27;
  DevBuf front;
  EDLGT_TRY(front.alloc(sizeof(uint64_t)));
  EDLGT_HIP(hipMemsetAsync(front.p, 0, sizeof(uint64_t), st));
  int64_t n_front = 1;
  int p = 0;
  while (p < n_sites && n_front > 0) {
    const int64_t n_cand = n_front * loc_dims[p];
    if (p > 0 && (n_front >= kTargetFrontier || n_cand > kMaxCandidates)) break;
    if (p > 0 && prodd / strides128[p] > (u128)(UINT64_MAX / 2))
      break;  // prefix indices of depth p+1 would not fit 64 bits
    DevBuf flags, pos, next;
    EDLGT_TRY(flags.alloc(n_cand * sizeof(int32_t)));
    EDLGT_TRY(pos.alloc((n_cand + 1) * sizeof(int64_t)));
    hipLaunchKernelGGL(expand_flags, dim3(grid_for(n_cand, 256)), dim3(256), 0, st, D,
                       front.as<uint64_t>(), n_front, p, flags.as<int32_t>());
    EDLGT_HIP(hipGetLastError());
    EDLGT_TRY(exclusive_scan_i32_i64(flags.as<int32_t>(), pos.as<int64_t>(), n_cand, st));
    int64_t n_next = 0;
    EDLGT_HIP(hipMemcpyAsync(&n_next, pos.as<int64_t>() + n_cand, sizeof(int64_t),
                             hipMemcpyDeviceToHost, st));
    EDLGT_HIP(hipStreamSynchronize(st));
    EDLGT_TRY(next.alloc(std::max<int64_t>(n_next, 1) * sizeof(uint64_t)));
    hipLaunchKernelGGL(expand_compact, dim3(grid_for(n_cand, 256)), dim3(256), 0, st,
                       front.as<uint64_t>(), n_front, (int64_t)loc_dims[p], flags.as<int32_t>(),
                       pos.as<int64_t>(), next.as<uint64_t>());
    EDLGT_HIP(hipGetLastError());
    EDLGT_HIP(hipStreamSynchronize(st));
    front = std::move(next);
    n_front = n_next;
    ++p;
  }

  out->n_sites = n_sites;
  out->cfg_stride = (n_sites + 3) & ~3;
  out->loc_dims.assign(loc_dims, loc_dims + n_sites);
  out->strides = strides;
  out->strides128 = strides128;
  out->key_words = wide ? 2 : 1;
  const size_t kb = wide ? 16 : 8;
  DevBuf b_str128;
  EDLGT_TRY(upload(b_str128, strides128.data(), strides128.size(), st));
  if (n_front == 0) {
    out->n_states = 0;
    EDLGT_TRY(out->keys.alloc(16));
    EDLGT_TRY(out->configs.alloc(8));
    return EDLGT_OK;
  }
  // Phase 2: depth-first completion, count -> scan -> fill.
  DevBuf counts, offs;
  EDLGT_TRY(counts.alloc(n_front * sizeof(int64_t)));
  EDLGT_TRY(offs.alloc((n_front + 1) * sizeof(int64_t)));
  const unsigned g2 = grid_for(n_front, 64);
  if (int_mode && wide)
    hipLaunchKernelGGL((complete_prefixes_int<false, u128>), dim3(g2), dim3(64), 0, st, D, IG,
                       b_str128.as<u128>(), front.as<uint64_t>(), n_front, p, counts.as<int64_t>(),
                       (const int64_t*)nullptr, (u128*)nullptr, (uint8_t*)nullptr, 0);
  else if (int_mode)
    hipLaunchKernelGGL((complete_prefixes_int<false, uint64_t>), dim3(g2), dim3(64), 0, st, D, IG,
                       b_str.as<uint64_t>(), front.as<uint64_t>(), n_front, p, counts.as<int64_t>(),
                       (const int64_t*)nullptr, (uint64_t*)nullptr, (uint8_t*)nullptr, 0);
  else
    hipLaunchKernelGGL(complete_prefixes<false>, dim3(g2), dim3(64), 0, st, D, front.as<uint64_t>(),
                       n_front, p, counts.as<int64_t>(), (const int64_t*)nullptr, (uint64_t*)nullptr,
                       (uint8_t*)nullptr, 0);
  EDLGT_HIP(hipGetLastError());
  EDLGT_TRY(exclusive_scan_i64(counts.as<int64_t>(), offs.as<int64_t>(), n_front, st));
  int64_t N = 0;
  EDLGT_HIP(hipMemcpyAsync(&N, offs.as<int64_t>() + n_front, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  out->n_states = N;
  EDLGT_TRY(out->keys.alloc(std::max<int64_t>(N, 1) * kb));
  EDLGT_TRY(out->configs.alloc(std::max<int64_t>(N, 1) * out->cfg_stride));
  EDLGT_HIP(hipMemsetAsync(out->configs.p, 0, std::max<int64_t>(N, 1) * out->cfg_stride, st));
  if (int_mode && wide)
    hipLaunchKernelGGL((complete_prefixes_int<true, u128>), dim3(g2), dim3(64), 0, st, D, IG,
                       b_str128.as<u128>(), front.as<uint64_t>(), n_front, p, (int64_t*)nullptr,
                       offs.as<int64_t>(), out->keys.as<u128>(), out->configs.as<uint8_t>(),
                       out->cfg_stride);
  else if (int_mode)
    hipLaunchKernelGGL((complete_prefixes_int<true, uint64_t>), dim3(g2), dim3(64), 0, st, D, IG,
                       b_str.as<uint64_t>(), front.as<uint64_t>(), n_front, p, (int64_t*)nullptr,
                       offs.as<int64_t>(), out->keys.as<uint64_t>(), out->configs.as<uint8_t>(),
                       out->cfg_stride);
  else
    hipLaunchKernelGGL(complete_prefixes<true>, dim3(g2), dim3(64), 0, st, D, front.as<uint64_t>(),
                       n_front, p, (int64_t*)nullptr, offs.as<int64_t>(), out->keys.as<uint64_t>(),
                       out->configs.as<uint8_t>(), out->cfg_stride);
  EDLGT_HIP(hipGetLastError());
  EDLGT_HIP(hipStreamSynchronize(st));
  return EDLGT_OK;
}

}  // namespace edlgt
