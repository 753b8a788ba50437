/*
 * refcpu -- C/OpenMP restatement of the reference's Numba CPU kernels for the
 * exact-diagonalization hot path.  TEST INFRASTRUCTURE / CPU BASELINE ONLY:
 * the shipped engine never links or calls this file.
 *
 * Restated kernels (edlgt 0.2.3):
 *   refcpu_csr_spmv_c128 / _f64   csr_spmv, edlgt/modeling/qmb_hamiltonian.py:705-730
 *                                 (prange over rows, y_i = sum_p data[p] x[idx[p]])
 *   refcpu_sector                 iterative_sitebased_sym_sector_configs /
 *                                 iterative_link_sector_configs_plus,
 *                                 edlgt/symmetries/symmetry_sector.py:785-967, 1329-1469
 *                                 (level-by-level prefix expansion with activated checks)
 *   refcpu_nbody_count / _fill    _count_row_nnz_from_transition_data + _get_nnz_offsets +
 *                                 nbody_data (edlgt/symmetries/sym_qmb_operations.py:95-396)
 *                                 with config_to_index_binarysearch (generate_configs.py:145-173)
 *
 * Build: oracle/Makefile.  librefcpu.so: gcc -O3 -march=x86-64-v3 -fopenmp -shared (AVX2 +
 * FMA, loadable on every x86-64 host of this image: the checker library is built once and
 * travels).  librefcpu_native.so (`make native`): the same with -march=native, built on the
 * host that runs the CPU baseline, which is timed with it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int refcpu_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* OpenMP team size of the following calls (1 = the one-thread baseline). */
void refcpu_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}

/* y = H x, complex128 values/vectors as interleaved (re, im) doubles. */
void refcpu_csr_spmv_c128(int64_t n_rows, const int64_t* indptr, const int32_t* indices,
                          const double* data, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n_rows; ++r) {
    double re = 0.0, im = 0.0;
    for (int64_t p = indptr[r]; p < indptr[r + 1]; ++p) {
      const double vr = data[2 * p], vi = data[2 * p + 1];
      const double xr = x[2 * (int64_t)indices[p]], xi = x[2 * (int64_t)indices[p] + 1];
      re += vr * xr - vi * xi;
      im += vr * xi + vi * xr;
    }
    y[2 * r] = re;
    y[2 * r + 1] = im;
  }
}

void refcpu_csr_spmv_f64(int64_t n_rows, const int64_t* indptr, const int32_t* indices,
                         const double* data, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n_rows; ++r) {
    double s = 0.0;
    for (int64_t p = indptr[r]; p < indptr[r + 1]; ++p) s += data[p] * x[indices[p]];
    y[r] = s;
  }
}

/* ------------------------------------------------------------------------
 * Sector enumeration: prefix expansion, one level per site, candidates
 * c = prev * d_k + s checked in parallel and compacted in order.
 * glob: (n_glob, n_sites, d_max) U(1) diagonals; link pairs (dir, i, j);
 * nbody groups (offsets/sites).  Returns the number of configurations and a
 * malloc'ed uint16 table in *out (caller frees with refcpu_free).
 * ---------------------------------------------------------------------- */
#define SYM_TOL 1e-13

static int check_level(int k, const uint16_t* cfg, int n_sites, int d_max, int n_glob,
                       const double* glob, const double* gsec, const double* gval_prev,
                       double* gval_new, const double* rmin, const double* rmax, int n_pairs,
                       const int32_t* pdir, const int32_t* p0, const int32_t* p1,
                       const double* link, const double* lsec, int n_nb, const int32_t* nboff,
                       const int32_t* nbsites, const double* nbd, const double* nbsec) {
  for (int g = 0; g < n_glob; ++g) {
    const double v = (k == 0 ? 0.0 : gval_prev[g]) + glob[((size_t)g * n_sites + k) * d_max + cfg[k]];
    gval_new[g] = v;
    if (k < n_sites - 1) {
      if (gsec[g] < v + rmin[g * (n_sites + 1) + k + 1] - SYM_TOL) return 0;
      if (gsec[g] > v + rmax[g * (n_sites + 1) + k + 1] + SYM_TOL) return 0;
    } else if (fabs(v - gsec[g]) > SYM_TOL) {
      return 0;
    }
  }
  for (int q = 0; q < n_pairs; ++q) {
    const int i = p0[q], j = p1[q];
    if ((i > j ? i : j) != k) continue;
    const int d = pdir[q];
    const double v = link[(((size_t)d * 2 + 0) * n_sites + i) * d_max + cfg[i]] +
                     link[(((size_t)d * 2 + 1) * n_sites + j) * d_max + cfg[j]];
    if (fabs(v - lsec[d]) > SYM_TOL) return 0;
  }
  for (int s = 0; s < n_nb; ++s) {
    int mx = -1;
    for (int q = nboff[s]; q < nboff[s + 1]; ++q) mx = nbsites[q] > mx ? nbsites[q] : mx;
    if (mx != k) continue;
    double v = 0.0;
    for (int q = nboff[s]; q < nboff[s + 1]; ++q)
      v += nbd[((size_t)s * n_sites + nbsites[q]) * d_max + cfg[nbsites[q]]];
    if (fabs(v - nbsec[s]) > SYM_TOL) return 0;
  }
  return 1;
}

int64_t refcpu_sector(int n_sites, const int32_t* loc_dims, int d_max, int n_glob,
                      const double* glob, const double* gsec, int n_pairs, const int32_t* pdir,
                      const int32_t* p0, const int32_t* p1, const double* link, const double* lsec,
                      int n_nb, const int32_t* nboff, const int32_t* nbsites, const double* nbd,
                      const double* nbsec, uint16_t** out) {
  double* rmin = calloc((size_t)(n_glob > 0 ? n_glob : 1) * (n_sites + 1), sizeof(double));
  double* rmax = calloc((size_t)(n_glob > 0 ? n_glob : 1) * (n_sites + 1), sizeof(double));
  for (int g = 0; g < n_glob; ++g)
    for (int s = n_sites - 1; s >= 0; --s) {
      const double* v = glob + ((size_t)g * n_sites + s) * d_max;
      double mn = v[0], mx = v[0];
      for (int a = 1; a < loc_dims[s]; ++a) {
        mn = v[a] < mn ? v[a] : mn;
        mx = v[a] > mx ? v[a] : mx;
      }
      rmin[g * (n_sites + 1) + s] = rmin[g * (n_sites + 1) + s + 1] + mn;
      rmax[g * (n_sites + 1) + s] = rmax[g * (n_sites + 1) + s + 1] + mx;
    }
  int64_t n_prev = 1;
  uint16_t* prev = calloc(1, sizeof(uint16_t));
  double* gprev = calloc((size_t)(n_glob > 0 ? n_glob : 1), sizeof(double));
  for (int k = 0; k < n_sites; ++k) {
    const int d = loc_dims[k];
    const int64_t n_cand = n_prev * d;
    uint16_t* cand = malloc((size_t)n_cand * (k + 1) * sizeof(uint16_t));
    double* gcand = malloc((size_t)n_cand * (n_glob > 0 ? n_glob : 1) * sizeof(double));
    char* ok = malloc((size_t)n_cand);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n_cand; ++c) {
      const int64_t p = c / d;
      uint16_t* row = cand + (size_t)c * (k + 1);
      if (k) memcpy(row, prev + (size_t)p * k, (size_t)k * sizeof(uint16_t));
      row[k] = (uint16_t)(c % d);
      ok[c] = (char)check_level(k, row, n_sites, d_max, n_glob, glob, gsec,
                                gprev + (size_t)p * (n_glob > 0 ? n_glob : 1),
                                gcand + (size_t)c * (n_glob > 0 ? n_glob : 1), rmin, rmax, n_pairs,
                                pdir, p0, p1, link, lsec, n_nb, nboff, nbsites, nbd, nbsec);
    }
    int64_t n_next = 0;
    for (int64_t c = 0; c < n_cand; ++c) {
      if (!ok[c]) continue;
      if (n_next != c) {
        memmove(cand + (size_t)n_next * (k + 1), cand + (size_t)c * (k + 1), (size_t)(k + 1) * sizeof(uint16_t));
        memmove(gcand + (size_t)n_next * (n_glob > 0 ? n_glob : 1), gcand + (size_t)c * (n_glob > 0 ? n_glob : 1),
                (size_t)(n_glob > 0 ? n_glob : 1) * sizeof(double));
      }
      ++n_next;
    }
    free(ok);
    free(prev);
    free(gprev);
    prev = cand;
    gprev = gcand;
    n_prev = n_next;
  }
  free(rmin);
  free(rmax);
  free(gprev);
  *out = prev;
  return n_prev;
}

void refcpu_free(void* p) { free(p); }

/* ------------------------------------------------------------------------
 * Two-pass n-body triplets (one term placement), binary search on rows.
 * counts[f*d_max + a], kets/vals[(f*d_max + a)*d_max + t]: transitions of
 * factor f from bra state a (|O| > 1e-10), vals complex interleaved.
 * ---------------------------------------------------------------------- */
static int64_t bsearch_row(const uint16_t* cfgs, int64_t n, int n_sites, const uint16_t* key) {
  int64_t lo = 0, hi = n - 1;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) >> 1;
    const uint16_t* row = cfgs + (size_t)mid * n_sites;
    int cmp = 0;
    for (int s = 0; s < n_sites; ++s) {
      if (row[s] != key[s]) {
        cmp = row[s] < key[s] ? -1 : 1;
        break;
      }
    }
    if (cmp == 0) return mid;
    if (cmp < 0) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

/* Pass 1: per-row counts and exclusive offsets (int64); returns total. */
int64_t refcpu_nbody_count(const uint16_t* cfgs, int64_t n, int n_sites, int n_ops,
                           const int32_t* sites, const int32_t* counts, int d_max, int64_t* offsets) {
  int64_t* c = malloc((size_t)n * sizeof(int64_t));
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    int64_t k = 1;
    for (int f = 0; f < n_ops; ++f) k *= counts[f * d_max + cfgs[(size_t)r * n_sites + sites[f]]];
    c[r] = k;
  }
  int64_t tot = 0;
  for (int64_t r = 0; r < n; ++r) {
    offsets[r] = tot;
    tot += c[r];
  }
  free(c);
  return tot;
}

/* Pass 2: triplets in row order, mixed-radix over the factors (last fastest).
 * Returns the number of lookup misses (written as -1, like the reference). */
int64_t refcpu_nbody_fill(const uint16_t* cfgs, int64_t n, int n_sites, int n_ops,
                          const int32_t* sites, const int32_t* counts, const int32_t* kets,
                          const double* vals, int d_max, const int64_t* offsets, int32_t* rows,
                          int32_t* cols, double* out_vals) {
  int64_t misses = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : misses)
  for (int64_t r = 0; r < n; ++r) {
    uint16_t ket[64];
    int a[16], cnt[16], idx[16];
    const uint16_t* bra = cfgs + (size_t)r * n_sites;
    memcpy(ket, bra, (size_t)n_sites * sizeof(uint16_t));
    int64_t w = offsets[r];
    int any = 1;
    for (int f = 0; f < n_ops; ++f) {
      a[f] = bra[sites[f]];
      cnt[f] = counts[f * d_max + a[f]];
      idx[f] = 0;
      if (cnt[f] == 0) any = 0;
    }
    if (!any) continue;
    for (;;) {
      double re = 1.0, im = 0.0;
      for (int f = 0; f < n_ops; ++f) {
        const size_t e = ((size_t)f * d_max + a[f]) * d_max + idx[f];
        ket[sites[f]] = (uint16_t)kets[e];
        const double vr = vals[2 * e], vi = vals[2 * e + 1];
        const double nr = re * vr - im * vi, ni = re * vi + im * vr;
        re = nr;
        im = ni;
      }
      const int64_t c = bsearch_row(cfgs, n, n_sites, ket);
      if (c < 0) ++misses;
      rows[w] = (int32_t)r;
      cols[w] = (int32_t)c;
      out_vals[2 * w] = re;
      out_vals[2 * w + 1] = im;
      ++w;
      int f = n_ops - 1;
      for (; f >= 0; --f) {
        if (++idx[f] < cnt[f]) break;
        idx[f] = 0;
      }
      if (f < 0) break;
    }
  }
  return misses;
}
