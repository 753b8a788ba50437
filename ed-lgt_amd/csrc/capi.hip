// C ABI of libedlgt_hip (declared in include/edlgt_hip.h): argument checks,
// error capture and handle management around the device kernels.
#include <cstring>
#include <memory>
#include <mutex>
#include <new>

#include "common.h"
#include "kernels.h"

namespace edlgt {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

namespace {

template <class K>
__device__ inline int64_t lb_key(const K* a, int64_t n, K k) {
  int64_t lo = 0, len = n;
  while (len > 0) {
    const int64_t half = len >> 1;
    if (a[lo + half] < k) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo;
}

template <class K>
__global__ void lookup_kernel(const K* __restrict__ keys, int64_t n, const K* __restrict__ q,
                              const uint8_t* __restrict__ bad, int64_t nq, int64_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = lb_key<K>(keys, n, q[i]);
    out[i] = (!bad[i] && p < n && keys[p] == q[i]) ? p : -1;
  }
}

}  // namespace

int basis_lookup(const Basis& b, const uint16_t* configs, int64_t nq, int64_t* out_idx) {
  if (nq <= 0) return EDLGT_OK;
  std::vector<u128> qk(nq);
  std::vector<uint8_t> bad(nq, 0);
  for (int64_t i = 0; i < nq; ++i) {
    u128 k = 0;
    for (int s = 0; s < b.n_sites; ++s) {
      const uint16_t c = configs[i * b.n_sites + s];
      if (c >= b.loc_dims[s]) bad[i] = 1;
      k += (u128)c * b.strides128[s];
    }
    qk[i] = k;
  }
  DevBuf dq, dbad, dout;
  EDLGT_TRY(dout.alloc(nq * sizeof(int64_t)));
  EDLGT_TRY(upload(dbad, bad.data(), nq, nullptr));
  if (b.key_words == 2) {
    EDLGT_TRY(upload(dq, qk.data(), nq, nullptr));
    hipLaunchKernelGGL(lookup_kernel<u128>, dim3(grid_for(nq, 256)), dim3(256), 0, nullptr,
                       b.keys.as<u128>(), b.n_states, dq.as<u128>(), dbad.as<uint8_t>(), nq,
                       dout.as<int64_t>());
  } else {
    std::vector<uint64_t> q64(nq);
    for (int64_t i = 0; i < nq; ++i) q64[i] = (uint64_t)qk[i];
    EDLGT_TRY(upload(dq, q64.data(), nq, nullptr));
    hipLaunchKernelGGL(lookup_kernel<uint64_t>, dim3(grid_for(nq, 256)), dim3(256), 0, nullptr,
                       b.keys.as<uint64_t>(), b.n_states, dq.as<uint64_t>(), dbad.as<uint8_t>(), nq,
                       dout.as<int64_t>());
  }
  EDLGT_HIP(hipGetLastError());
  EDLGT_HIP(hipMemcpy(out_idx, dout.p, nq * sizeof(int64_t), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < nq; ++i)
    if (out_idx[i] < 0)
      return set_error(EDLGT_ERR_KEY_NOT_FOUND, "config not compatible with the symmetry sector");
  return EDLGT_OK;
}

}  // namespace edlgt

using namespace edlgt;

#define EDLGT_REQUIRE(cond, msg) \
  do {                           \
    if (!(cond)) return set_error(EDLGT_ERR_INVALID_ARG, msg); \
  } while (0)

extern "C" {

int edlgt_last_error(char* buf, int64_t size) {
  if (!buf || size <= 0) return EDLGT_ERR_INVALID_ARG;
  const size_t n = std::min<size_t>((size_t)size - 1, g_last_error.size());
  std::memcpy(buf, g_last_error.data(), n);
  buf[n] = '\0';
  return EDLGT_OK;
}

int edlgt_version(void) { return 1; }

int edlgt_device_info(int32_t* device, int64_t* free_bytes, int64_t* total_bytes) {
  int d = 0;
  EDLGT_HIP(hipGetDevice(&d));
  size_t fr = 0, tot = 0;
  EDLGT_HIP(hipMemGetInfo(&fr, &tot));
  if (device) *device = d;
  if (free_bytes) *free_bytes = (int64_t)fr;
  if (total_bytes) *total_bytes = (int64_t)tot;
  return EDLGT_OK;
}

int edlgt_sector_build(int32_t n_sites, const int32_t* loc_dims, int32_t d_max, int32_t n_glob,
                       int32_t glob_type, const double* glob_diags, const double* glob_sectors,
                       int32_t n_dirs, const double* link_diags, const double* link_sectors,
                       int32_t n_pairs, const int32_t* pair_dir, const int32_t* pair_site0,
                       const int32_t* pair_site1, int32_t n_nbody, int32_t nbody_type,
                       const double* nbody_diags, const double* nbody_sectors,
                       const int32_t* nbody_site_offsets, const int32_t* nbody_sites,
                       void* stream, edlgt_basis** out) {
  EDLGT_REQUIRE(out && loc_dims, "null output or loc_dims");
  *out = nullptr;
  EDLGT_REQUIRE(n_glob == 0 || (glob_diags && glob_sectors), "null global tables");
  EDLGT_REQUIRE(n_pairs == 0 || (pair_dir && pair_site0 && pair_site1 && link_diags && link_sectors),
                "null link tables");
  EDLGT_REQUIRE(n_nbody == 0 || (nbody_diags && nbody_sectors && nbody_site_offsets && nbody_sites),
                "null n-body tables");
  EDLGT_REQUIRE(n_nbody == 0 || nbody_type == EDLGT_SYM_U1 || nbody_type == EDLGT_SYM_Z2 ||
                    nbody_type >= 18,
                "nbody_type must be U1, Z2 or EDLGT_SYM_MOD(n >= 2)");
  edlgt_basis* b = new (std::nothrow) edlgt_basis();
  if (!b) return set_error(EDLGT_ERR_OUT_OF_MEMORY, "host allocation failed");
  int s = sector_build(n_sites, loc_dims, d_max, n_glob, glob_type, glob_diags, glob_sectors, n_dirs,
                       link_diags, link_sectors, n_pairs, pair_dir, pair_site0, pair_site1, n_nbody,
                       nbody_type, nbody_diags, nbody_sectors, nbody_site_offsets, nbody_sites,
                       (hipStream_t)stream, b);
  if (s != EDLGT_OK) {
    delete b;
    return s;
  }
  *out = b;
  return EDLGT_OK;
}

int edlgt_basis_info(const edlgt_basis* b, int64_t* n_states, int32_t* n_sites) {
  EDLGT_REQUIRE(b, "null basis");
  if (n_states) *n_states = b->n_states;
  if (n_sites) *n_sites = b->n_sites;
  return EDLGT_OK;
}

int edlgt_basis_key_words(const edlgt_basis* b, int32_t* key_words) {
  EDLGT_REQUIRE(b && key_words, "null argument");
  *key_words = b->key_words;
  return EDLGT_OK;
}

int edlgt_basis_copy_configs(const edlgt_basis* b, uint16_t* out_host) {
  EDLGT_REQUIRE(b && out_host, "null argument");
  const int64_t N = b->n_states;
  if (N == 0) return EDLGT_OK;
  std::vector<uint8_t> raw((size_t)N * b->cfg_stride);
  EDLGT_HIP(hipMemcpy(raw.data(), b->configs.p, raw.size(), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < N; ++i)
    for (int s = 0; s < b->n_sites; ++s) out_host[i * b->n_sites + s] = raw[i * b->cfg_stride + s];
  return EDLGT_OK;
}

int edlgt_basis_copy_keys(const edlgt_basis* b, uint64_t* out_host) {
  EDLGT_REQUIRE(b && out_host, "null argument");
  if (b->n_states == 0) return EDLGT_OK;
  EDLGT_HIP(hipMemcpy(out_host, b->keys.p, b->n_states * 8 * b->key_words, hipMemcpyDeviceToHost));
  return EDLGT_OK;
}

int edlgt_basis_lookup(const edlgt_basis* b, const uint16_t* configs_host, int64_t n_query,
                       int64_t* out_idx_host) {
  EDLGT_REQUIRE(b && (n_query == 0 || (configs_host && out_idx_host)), "null argument");
  return basis_lookup(*b, configs_host, n_query, out_idx_host);
}

int edlgt_basis_device_ptrs(const edlgt_basis* b, void** keys, void** configs, int32_t* cfg_stride) {
  EDLGT_REQUIRE(b, "null basis");
  if (keys) *keys = b->keys.p;
  if (configs) *configs = b->configs.p;
  if (cfg_stride) *cfg_stride = b->cfg_stride;
  return EDLGT_OK;
}

int edlgt_basis_destroy(edlgt_basis* b) {
  delete b;
  return EDLGT_OK;
}

int edlgt_expval_terms(const edlgt_basis* b, int32_t d_max, int32_t n_tables, const double* tables,
                       int32_t n_place, const int32_t* place_nops, const int32_t* place_sites,
                       const int32_t* place_tables, const double* place_strength, const void* psi,
                       int32_t vec_type, double* out_host, void* stream) {
  EDLGT_REQUIRE(b && tables && psi && (out_host || n_place == 0), "null argument");
  EDLGT_REQUIRE(n_place == 0 || (place_nops && place_sites && place_tables && place_strength),
                "null placement arrays");
  return placements_expval(*b, d_max, n_tables, tables, n_place, place_nops, place_sites, place_tables,
                           place_strength, psi, vec_type, out_host, (hipStream_t)stream);
}

int edlgt_hamiltonian_count(const edlgt_basis* b, int32_t d_max, int32_t n_tables, const double* tables,
                            int32_t n_place, const int32_t* place_nops, const int32_t* place_sites,
                            const int32_t* place_tables, const double* place_strength,
                            const int32_t* place_dagger, void* stream, int64_t* out_nnz) {
  EDLGT_REQUIRE(b && out_nnz && tables, "null argument");
  EDLGT_REQUIRE(n_place == 0 || (place_nops && place_sites && place_tables && place_strength && place_dagger),
                "null placement arrays");
  return hamiltonian_count(*b, d_max, n_tables, tables, n_place, place_nops, place_sites, place_tables,
                           place_strength, place_dagger, (hipStream_t)stream, out_nnz);
}

int edlgt_hamiltonian_build_matfree(const edlgt_basis* b, int32_t d_max, int32_t n_tables,
                                    const double* tables, int32_t n_place, const int32_t* place_nops,
                                    const int32_t* place_sites, const int32_t* place_tables,
                                    const double* place_strength, const int32_t* place_dagger,
                                    int32_t value_type, void* stream, edlgt_csr** out) {
  EDLGT_REQUIRE(b && out && tables, "null argument");
  EDLGT_REQUIRE(n_place == 0 || (place_nops && place_sites && place_tables && place_strength && place_dagger),
                "null placement arrays");
  *out = nullptr;
  auto h = std::make_unique<edlgt_csr>();
  EDLGT_TRY(hamiltonian_build_matfree(*b, d_max, n_tables, tables, n_place, place_nops, place_sites, place_tables,
                                      place_strength, place_dagger, value_type, (hipStream_t)stream, h.get()));
  *out = h.release();
  return EDLGT_OK;
}

int edlgt_csr_is_matrix_free(const edlgt_csr* h, int32_t* out) {
  EDLGT_REQUIRE(h && out, "null argument");
  *out = h->mf ? 1 : 0;
  return EDLGT_OK;
}

int edlgt_hamiltonian_build(const edlgt_basis* b, int32_t d_max, int32_t n_tables,
                            const double* tables, int32_t n_place, const int32_t* place_nops,
                            const int32_t* place_sites, const int32_t* place_tables,
                            const double* place_strength, const int32_t* place_dagger,
                            int32_t value_type, void* stream, edlgt_csr** out) {
  EDLGT_REQUIRE(b && out && tables, "null argument");
  EDLGT_REQUIRE(n_place == 0 || (place_nops && place_sites && place_tables && place_strength &&
                                 place_dagger),
                "null placement arrays");
  *out = nullptr;
  edlgt_csr* h = new (std::nothrow) edlgt_csr();
  if (!h) return set_error(EDLGT_ERR_OUT_OF_MEMORY, "host allocation failed");
  int s = hamiltonian_build(*b, d_max, n_tables, tables, n_place, place_nops, place_sites,
                            place_tables, place_strength, place_dagger, nullptr, 1, value_type,
                            (hipStream_t)stream, h);
  if (s != EDLGT_OK) {
    delete h;
    return s;
  }
  *out = h;
  return EDLGT_OK;
}

int edlgt_hamiltonian_build_terms(const edlgt_basis* b, int32_t d_max, int32_t n_tables,
                                  const double* tables, int32_t n_place, const int32_t* place_nops,
                                  const int32_t* place_sites, const int32_t* place_tables,
                                  const double* place_strength, const int32_t* place_dagger,
                                  const int32_t* place_term, int32_t n_terms, int32_t value_type,
                                  void* stream, edlgt_csr** out) {
  EDLGT_REQUIRE(b && out && tables && place_term, "null argument");
  EDLGT_REQUIRE(n_place == 0 || (place_nops && place_sites && place_tables && place_strength &&
                                 place_dagger),
                "null placement arrays");
  *out = nullptr;
  edlgt_csr* h = new (std::nothrow) edlgt_csr();
  if (!h) return set_error(EDLGT_ERR_OUT_OF_MEMORY, "host allocation failed");
  int s = hamiltonian_build(*b, d_max, n_tables, tables, n_place, place_nops, place_sites,
                            place_tables, place_strength, place_dagger, place_term, n_terms,
                            value_type, (hipStream_t)stream, h);
  if (s != EDLGT_OK) {
    delete h;
    return s;
  }
  *out = h;
  return EDLGT_OK;
}

int edlgt_csr_copy_terms(const edlgt_csr* h, uint8_t* term_host, int32_t* n_terms) {
  EDLGT_REQUIRE(h, "null matrix");
  EDLGT_REQUIRE(!h->mf, "not available for a matrix-free operator (no stored entries)");
  EDLGT_REQUIRE(!term_host || h->n_terms == 1 || h->term.p, "the term array was released (coded matrix)");
  if (n_terms) *n_terms = h->n_terms;
  if (term_host && h->n_terms > 1 && h->nnz)
    EDLGT_HIP(hipMemcpy(term_host, h->term.p, h->nnz, hipMemcpyDeviceToHost));
  return EDLGT_OK;
}

int edlgt_csr_compress(edlgt_csr* h, void* stream, int32_t* n_dict, int32_t* code_bytes) {
  EDLGT_REQUIRE(h, "null matrix");
  EDLGT_REQUIRE(!h->mf, "not available for a matrix-free operator (no stored entries)");
  int done = 0;
  EDLGT_TRY(csr_compress(*h, (hipStream_t)stream, &done));
  if (n_dict) *n_dict = h->n_dict;
  if (code_bytes) *code_bytes = h->code_bytes;
  return EDLGT_OK;
}

int edlgt_csr_release_values(edlgt_csr* h) {
  EDLGT_REQUIRE(h, "null matrix");
  EDLGT_REQUIRE(!h->mf, "not available for a matrix-free operator (no stored entries)");
  EDLGT_REQUIRE(h->n_dict > 0, "only a coded matrix can release its value array");
  EDLGT_HIP(hipDeviceSynchronize());
  h->val.release();
  h->val_lo.release();
  h->term.release();
  return EDLGT_OK;
}

int edlgt_csr_info(const edlgt_csr* h, int64_t* n_rows, int64_t* nnz, int32_t* value_type) {
  EDLGT_REQUIRE(h, "null matrix");
  if (n_rows) *n_rows = h->n_rows;
  if (nnz) *nnz = h->nnz;
  if (value_type) *value_type = h->value_type;
  return EDLGT_OK;
}

int edlgt_csr_copy(const edlgt_csr* h, int64_t* row_ptr, int32_t* col, void* val) {
  EDLGT_REQUIRE(h, "null matrix");
  EDLGT_REQUIRE(!h->mf, "not available for a matrix-free operator (no stored entries)");
  EDLGT_REQUIRE(!val || h->val.p || h->nnz == 0, "the value array was released (coded matrix)");
  if (row_ptr)
    EDLGT_HIP(hipMemcpy(row_ptr, h->row_ptr.p, (h->n_rows + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (col && h->nnz)
    EDLGT_HIP(hipMemcpy(col, h->col.p, h->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (val && h->nnz)
    EDLGT_HIP(hipMemcpy(val, h->val.p, h->nnz * value_bytes(h->value_type), hipMemcpyDeviceToHost));
  return EDLGT_OK;
}

int edlgt_csr_device_ptrs(const edlgt_csr* h, void** row_ptr, void** col, void** val) {
  EDLGT_REQUIRE(h, "null matrix");
  EDLGT_REQUIRE(!h->mf, "not available for a matrix-free operator (no stored entries)");
  if (row_ptr) *row_ptr = h->row_ptr.p;
  if (col) *col = h->col.p;
  if (val) *val = h->val.p;
  return EDLGT_OK;
}

int edlgt_csr_dense_eigh(const edlgt_csr* h, double* evals, void* evecs, void* stream) {
  EDLGT_REQUIRE(h, "null matrix");
  return csr_dense_eigh(*h, evals, evecs, (hipStream_t)stream);
}

int edlgt_csr_make_low_precision(edlgt_csr* h, void* stream) {
  EDLGT_REQUIRE(h, "null matrix");
  return csr_make_low_precision(*h, (hipStream_t)stream);
}

int edlgt_csr_destroy(edlgt_csr* h) {
  delete h;
  return EDLGT_OK;
}

int edlgt_spmv(edlgt_csr* h, int32_t precision, const void* x, void* y, int32_t vec_type,
               void* stream) {
  return edlgt_spmm(h, precision, 1, x, y, vec_type, stream);
}

int edlgt_spmm(edlgt_csr* h, int32_t precision, int32_t n_rhs, const void* X, void* Y,
               int32_t vec_type, void* stream) {
  EDLGT_REQUIRE(h && X && Y, "null argument");
  EDLGT_REQUIRE(n_rhs >= 1, "n_rhs must be >= 1");
  EDLGT_REQUIRE(precision == EDLGT_PREC_FP64 || precision == EDLGT_PREC_MIXED, "bad precision");
  const bool lo_vec = vec_type == EDLGT_VALUE_C64 || vec_type == EDLGT_VALUE_F32;
  EDLGT_REQUIRE(vec_type == EDLGT_VALUE_C128 || vec_type == EDLGT_VALUE_F64 ||
                    (lo_vec && precision == EDLGT_PREC_MIXED),
                "vec_type must be C128 or F64 (or C64/F32 with EDLGT_PREC_MIXED)");
  EDLGT_REQUIRE(!((vec_type == EDLGT_VALUE_F64 || vec_type == EDLGT_VALUE_F32) &&
                  h->value_type != EDLGT_VALUE_F64),
                "real vectors need a real matrix");
  EDLGT_REQUIRE(!(vec_type == EDLGT_VALUE_C64 && h->value_type != EDLGT_VALUE_C128),
                "complex64 vectors need a complex matrix");
  if (precision == EDLGT_PREC_MIXED) EDLGT_TRY(csr_make_low_precision(*h, (hipStream_t)stream));
  if (n_rhs > 1 && precision == EDLGT_PREC_FP64 && spmm_packable(*h, vec_type)) {
    // row-interleaved groups of columns in a temporary workspace; the call
    // returns once the product is done (the workspace is released)
    DevBuf pack;
    EDLGT_TRY(pack.alloc(spmm_pack_bytes(*h, vec_type)));
    EDLGT_TRY(csr_spmm(*h, precision, vec_type, n_rhs, X, h->n_rows, Y, h->n_rows, (hipStream_t)stream, pack.p,
                       pack.bytes));
    EDLGT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return EDLGT_OK;
  }
  return csr_spmm(*h, precision, vec_type, n_rhs, X, h->n_rows, Y, h->n_rows, (hipStream_t)stream);
}

int edlgt_spmm_coef(edlgt_csr* h, int32_t n_rhs, const void* X, void* Y, int32_t vec_type,
                    const double* coef_host, void* stream) {
  EDLGT_REQUIRE(h && X && Y && coef_host, "null argument");
  EDLGT_REQUIRE(!h->mf, "not available for a matrix-free operator (no stored entries)");
  EDLGT_REQUIRE(n_rhs >= 1, "n_rhs must be >= 1");
  EDLGT_REQUIRE(vec_type == EDLGT_VALUE_C128 || vec_type == EDLGT_VALUE_F64, "vec_type must be C128 or F64");
  EDLGT_REQUIRE(!(vec_type == EDLGT_VALUE_F64 && h->value_type != EDLGT_VALUE_F64),
                "real vectors need a real matrix");
  hipStream_t st = (hipStream_t)stream;
  const int nt = h->n_terms;
  const size_t es = vec_type == EDLGT_VALUE_C128 ? 16 : 8;
  DevBuf xs, ys, cf, pack;
  EDLGT_TRY(xs.alloc(kCoefMaxRhs * sizeof(void*)));
  EDLGT_TRY(ys.alloc(kCoefMaxRhs * sizeof(void*)));
  EDLGT_TRY(cf.alloc(kCoefMaxRhs * 16 * sizeof(double)));
  EDLGT_TRY(pack.alloc(coef_pack_bytes(*h, vec_type, std::min(n_rhs, kCoefMaxRhs))));
  std::vector<const void*> hx(kCoefMaxRhs);
  std::vector<void*> hy(kCoefMaxRhs);
  std::vector<double> hc(kCoefMaxRhs * 16, 0.0);
  int done = 0;
  while (done < n_rhs) {
    const int nb = std::min(n_rhs - done, kCoefMaxRhs);
    for (int c = 0; c < nb; ++c) {
      hx[c] = (const char*)X + (size_t)(done + c) * h->n_rows * es;
      hy[c] = (char*)Y + (size_t)(done + c) * h->n_rows * es;
      for (int t = 0; t < 16; ++t) hc[c * 16 + t] = t < nt ? coef_host[(size_t)(done + c) * nt + t] : 0.0;
    }
    EDLGT_HIP(hipMemcpyAsync(xs.p, hx.data(), nb * sizeof(void*), hipMemcpyHostToDevice, st));
    EDLGT_HIP(hipMemcpyAsync(ys.p, hy.data(), nb * sizeof(void*), hipMemcpyHostToDevice, st));
    EDLGT_HIP(hipMemcpyAsync(cf.p, hc.data(), kCoefMaxRhs * 16 * sizeof(double), hipMemcpyHostToDevice, st));
    EDLGT_TRY(csr_spmm_coef(*h, vec_type, nb, xs.p, ys.p, hx.data(), hy.data(), cf.as<double>(), pack.p, st));
    EDLGT_HIP(hipStreamSynchronize(st));
    done += nb;
  }
  return EDLGT_OK;
}

int edlgt_csr_autotune(edlgt_csr* h, int32_t precision, int32_t vec_type, int32_t n_rhs,
                       const void* X, void* Y, void* stream, int32_t* out_lanes, int64_t* out_grid,
                       float* out_ms) {
  EDLGT_REQUIRE(h && X && Y && n_rhs >= 1, "null argument");
  EDLGT_REQUIRE(!h->mf, "not available for a matrix-free operator (no stored entries)");
  EDLGT_REQUIRE(precision == EDLGT_PREC_FP64 || precision == EDLGT_PREC_MIXED, "bad precision");
  if (precision == EDLGT_PREC_MIXED) EDLGT_TRY(csr_make_low_precision(*h, (hipStream_t)stream));
  return csr_autotune(*h, precision, vec_type, n_rhs, X, Y, (hipStream_t)stream, out_lanes, out_grid,
                      out_ms);
}

int edlgt_csr_set_launch(edlgt_csr* h, int32_t lanes_per_row, int64_t grid) {
  EDLGT_REQUIRE(h, "null matrix");
  EDLGT_REQUIRE(!h->mf, "not available for a matrix-free operator (no stored entries)");
  EDLGT_REQUIRE(lanes_per_row == 0 || lanes_per_row == 4 || lanes_per_row == 8 ||
                    lanes_per_row == 16 || lanes_per_row == 32 || lanes_per_row == 64,
                "lanes_per_row must be 0, 4, 8, 16, 32 or 64");
  if (grid < 0) {
    // grid = -(variant + 1) selects a launch variant of the stream kernel (the
    // coded one for a coded matrix); lanes_per_row must be 0
    EDLGT_REQUIRE(lanes_per_row == 0 && -grid - 1 < 5, "bad stream launch variant");
    const int v = (int)(-grid - 1);
    const int tiles[5] = {2048, 1024, 4096, 4096, 2048};
    EDLGT_TRY(csr_prepare_tiles(*h, tiles[v], nullptr));
    if (h->n_dict > 0)
      h->coded_variant = v;
    else
      h->plain_variant = v;
    h->g_override = 0;
    h->grid_override = 0;
    return EDLGT_OK;
  }
  h->g_override = lanes_per_row;
  h->grid_override = grid;
  return EDLGT_OK;
}

int edlgt_lanczos(edlgt_csr* h, int32_t n_eigs, double tol, int32_t max_basis, int64_t max_iter,
                  int32_t precision, int32_t vec_type, const void* v0, double* out_evals,
                  void* out_evecs, double* out_resid, int64_t* out_matvecs, void* stream) {
  EDLGT_REQUIRE(h && out_evals, "null argument");
  EDLGT_REQUIRE(tol > 0.0, "tol must be positive");
  EDLGT_REQUIRE(precision == EDLGT_PREC_FP64 || precision == EDLGT_PREC_MIXED, "bad precision");
  return lanczos_run(*h, n_eigs, tol, max_basis, max_iter, precision, vec_type, v0, out_evals,
                     out_evecs, out_resid, out_matvecs, (hipStream_t)stream);
}

int edlgt_basis_marginals(const edlgt_basis* b, const void* psi, int32_t vec_type, int32_t d_max,
                          double* out_host, void* stream) {
  EDLGT_REQUIRE(b && psi && out_host, "null argument");
  EDLGT_REQUIRE(vec_type == EDLGT_VALUE_C128 || vec_type == EDLGT_VALUE_F64, "vec_type must be C128 or F64");
  const int bins = b->n_sites * d_max;
  EDLGT_REQUIRE(d_max >= 1 && bins <= 4096, "n_sites * d_max must be <= 4096");
  DevBuf out;
  EDLGT_TRY(out.alloc((size_t)bins * sizeof(double)));
  EDLGT_TRY(basis_marginals(*b, psi, vec_type == EDLGT_VALUE_C128, d_max, out.as<double>(),
                            (hipStream_t)stream));
  EDLGT_HIP(hipMemcpy(out_host, out.p, (size_t)bins * sizeof(double), hipMemcpyDeviceToHost));
  return EDLGT_OK;
}

int edlgt_basis_marginals_noalloc(const edlgt_basis* b, const void* psi, int32_t vec_type,
                                  int32_t d_max, double* out_dev, void* stream) {
  EDLGT_REQUIRE(b && psi && out_dev, "null argument");
  EDLGT_REQUIRE(vec_type == EDLGT_VALUE_C128 || vec_type == EDLGT_VALUE_F64, "vec_type must be C128 or F64");
  EDLGT_REQUIRE(d_max >= 1 && b->n_sites * d_max <= 4096, "n_sites * d_max must be <= 4096");
  return basis_marginals_noalloc(*b, psi, vec_type == EDLGT_VALUE_C128, d_max, out_dev,
                                 (hipStream_t)stream);
}

int edlgt_basis_signed_image(const edlgt_basis* b, const int32_t* site_perm, const int32_t* loc_perm,
                             const double* loc_sign, int32_t d_max, int64_t* out_index, double* out_sign,
                             void* stream) {
  EDLGT_REQUIRE(b && site_perm && loc_perm && loc_sign, "null argument");
  EDLGT_REQUIRE(b->n_states == 0 || (out_index && out_sign), "null output");
  return basis_signed_image(*b, site_perm, loc_perm, loc_sign, d_max, out_index, out_sign, (hipStream_t)stream);
}

int edlgt_ensemble_density(const void* rows_matrix, int64_t n, int64_t n_rows, int64_t n_members,
                           const int64_t* member_rows, const double* member_coef, const int64_t* block_start,
                           int64_t n_blocks, double* density, void* stream) {
  return ensemble_density(rows_matrix, n, n_rows, n_members, member_rows, member_coef, block_start, n_blocks,
                          density, (hipStream_t)stream);
}

int edlgt_basis_sre2(const edlgt_basis* b, const void* psi, double prob_threshold, double* sre2,
                     int64_t* support_size, double* discarded_weight, int64_t* n_strings, void* stream) {
  EDLGT_REQUIRE(b && psi, "null argument");
  EDLGT_REQUIRE(sre2 && support_size && discarded_weight && n_strings, "null output");
  return basis_sre2(*b, psi, prob_threshold, sre2, support_size, discarded_weight, n_strings,
                    (hipStream_t)stream);
}

int edlgt_basis_quarter_phase(const edlgt_basis* b, const int32_t* site_exp, int32_t d_max, int32_t n_vec,
                              const double* x, void* y, void* stream) {
  EDLGT_REQUIRE(b && site_exp, "null argument");
  EDLGT_REQUIRE(n_vec >= 0, "negative vector count");
  EDLGT_REQUIRE(n_vec == 0 || b->n_states == 0 || (x && y), "null vector");
  EDLGT_REQUIRE(d_max >= 1 && b->n_sites * d_max <= 4096, "n_sites * d_max must be <= 4096");
  return basis_quarter_phase(*b, site_exp, d_max, n_vec, 1, false, x, y, (hipStream_t)stream);
}

int edlgt_basis_quarter_phase_c(const edlgt_basis* b, const int32_t* site_exp, int32_t d_max, int32_t n_vec,
                                int32_t direction, const void* x, void* y, void* stream) {
  EDLGT_REQUIRE(b && site_exp, "null argument");
  EDLGT_REQUIRE(n_vec >= 0, "negative vector count");
  EDLGT_REQUIRE(n_vec == 0 || b->n_states == 0 || (x && y), "null vector");
  EDLGT_REQUIRE(d_max >= 1 && b->n_sites * d_max <= 4096, "n_sites * d_max must be <= 4096");
  return basis_quarter_phase(*b, site_exp, d_max, n_vec, direction, true, x, y, (hipStream_t)stream);
}

int edlgt_schmidt_gram(const void* psi, int32_t vec_type, int64_t n_states, const int64_t* pos, int32_t n_blocks,
                       const int64_t* blocks, void* gram, int64_t gram_elems, void* stream) {
  EDLGT_REQUIRE(n_states >= 0 && n_blocks >= 0 && gram_elems >= 0, "negative size");
  EDLGT_REQUIRE(n_states == 0 || (psi && pos), "null psi or pos");
  EDLGT_REQUIRE(n_blocks == 0 || (blocks && gram), "null blocks or gram");
  EDLGT_REQUIRE(vec_type == EDLGT_VALUE_C128 || vec_type == EDLGT_VALUE_F64, "vec_type must be C128 or F64");
  return schmidt_gram(psi, vec_type == EDLGT_VALUE_C128, n_states, pos, n_blocks, blocks, gram, gram_elems,
                      (hipStream_t)stream);
}

int edlgt_hermitian_eigvals(void* gram, int64_t gram_elems, int32_t n_blocks, const int64_t* blocks, double* evals,
                            int64_t n_evals, void* stream) {
  EDLGT_REQUIRE(n_blocks >= 0 && gram_elems >= 0 && n_evals >= 0, "negative size");
  EDLGT_REQUIRE(n_blocks == 0 || (gram && blocks && evals), "null argument");
  return hermitian_eigvals(gram, gram_elems, n_blocks, blocks, evals, n_evals, (hipStream_t)stream);
}

int edlgt_random_uniform(void* out, int64_t n, int32_t value_type, const uint64_t* state, void* stream) {
  EDLGT_REQUIRE(n >= 0, "n must be >= 0");
  EDLGT_REQUIRE(n == 0 || (out && state), "null argument");
  EDLGT_REQUIRE(value_type == EDLGT_VALUE_F64 || value_type == EDLGT_VALUE_C128, "value_type must be F64 or C128");
  return random_uniform(out, n, value_type == EDLGT_VALUE_C128, state, (hipStream_t)stream);
}

int edlgt_krylov_evolve(edlgt_csr* h, const edlgt_basis* b, int32_t precision, void* psi, double dt,
                        int64_t n_steps, int32_t m, int32_t reorth, int32_t d_max,
                        double* out_marginals, double* out_err, void* psi_history, void* stream) {
  EDLGT_REQUIRE(h && psi, "null argument");
  EDLGT_REQUIRE(n_steps >= 0 && m >= 1, "n_steps must be >= 0 and m >= 1");
  EDLGT_REQUIRE(precision == EDLGT_PREC_FP64 || precision == EDLGT_PREC_MIXED, "bad precision");
  EDLGT_REQUIRE(!out_marginals || (b && d_max >= 1 && b->n_sites * d_max <= 4096),
                "marginals need the basis and n_sites * d_max <= 4096");
  EDLGT_REQUIRE(!b || b->n_states == h->n_rows, "basis and matrix sizes differ");
  return krylov_run(*h, b, precision, psi, dt, n_steps, m, reorth, d_max, out_marginals, out_err,
                    psi_history, (hipStream_t)stream);
}

int edlgt_momentum_build(const edlgt_basis* b, int32_t n_dims, const int32_t* lvals, const int32_t* unit_cell,
                         const int32_t* k_vals, void* stream, edlgt_momentum** out) {
  EDLGT_REQUIRE(b && lvals && unit_cell && k_vals && out, "null argument");
  *out = nullptr;
  auto m = std::make_unique<edlgt_momentum>();
  EDLGT_TRY(momentum_build(*b, n_dims, lvals, unit_cell, k_vals, (hipStream_t)stream, m.get()));
  *out = m.release();
  return EDLGT_OK;
}

int edlgt_momentum_build_tc(const edlgt_basis* b, int32_t d_loc, const int32_t* c_map, const double* c_sign,
                            int32_t k, void* stream, edlgt_momentum** out) {
  EDLGT_REQUIRE(b && c_map && c_sign && out, "null argument");
  *out = nullptr;
  auto m = std::make_unique<edlgt_momentum>();
  EDLGT_TRY(momentum_build_tc(*b, d_loc, c_map, c_sign, k, (hipStream_t)stream, m.get()));
  *out = m.release();
  return EDLGT_OK;
}

int edlgt_momentum_info(const edlgt_momentum* m, int64_t* n_rows, int64_t* n_cols, int64_t* nnz,
                        int32_t* is_complex) {
  EDLGT_REQUIRE(m, "null argument");
  if (n_rows) *n_rows = m->n_rows;
  if (n_cols) *n_cols = m->n_cols;
  if (nnz) *nnz = m->nnz;
  if (is_complex) *is_complex = m->is_complex;
  return EDLGT_OK;
}

int edlgt_momentum_copy(const edlgt_momentum* m, int64_t* col_ptr, int32_t* row_idx, double* data) {
  EDLGT_REQUIRE(m && col_ptr && row_idx && data, "null argument");
  return momentum_copy(*m, col_ptr, row_idx, data);
}

int edlgt_momentum_apply(const edlgt_momentum* m, int32_t forward, const void* in, void* out, void* stream) {
  EDLGT_REQUIRE(m && in && out, "null argument");
  return momentum_apply(*m, forward ? 1 : 0, in, out, (hipStream_t)stream);
}

int edlgt_csr_create(int64_t n, int64_t nnz, const int64_t* row_ptr, const int32_t* col, const void* val,
                     int32_t value_type, void* stream, edlgt_csr** out) {
  EDLGT_REQUIRE(out, "null output");
  *out = nullptr;
  EDLGT_REQUIRE(n >= 0 && nnz >= 0, "negative size");
  EDLGT_REQUIRE(row_ptr && (nnz == 0 || (col && val)), "null argument");
  EDLGT_REQUIRE(value_type == EDLGT_VALUE_C128 || value_type == EDLGT_VALUE_F64,
                "value_type must be EDLGT_VALUE_C128 or EDLGT_VALUE_F64");
  EDLGT_REQUIRE(n < 0x7fffffffLL, "row count above int32 column indices");
  EDLGT_REQUIRE(row_ptr[0] == 0 && row_ptr[n] == nnz, "row_ptr must run from 0 to nnz");
  for (int64_t r = 0; r < n; ++r)
    EDLGT_REQUIRE(row_ptr[r + 1] >= row_ptr[r], "row_ptr must not decrease");
  for (int64_t p = 0; p < nnz; ++p)
    EDLGT_REQUIRE(col[p] >= 0 && col[p] < n, "column index outside the matrix");
  hipStream_t st = (hipStream_t)stream;
  auto c = std::make_unique<edlgt_csr>();
  c->n_rows = n;
  c->nnz = nnz;
  c->value_type = value_type;
  EDLGT_TRY(upload(c->row_ptr, row_ptr, (size_t)n + 1, st));
  EDLGT_TRY(upload(c->col, col, (size_t)nnz, st));
  if (value_type == EDLGT_VALUE_C128)
    EDLGT_TRY(upload(c->val, static_cast<const double2*>(val), (size_t)nnz, st));
  else
    EDLGT_TRY(upload(c->val, static_cast<const double*>(val), (size_t)nnz, st));
  EDLGT_TRY(csr_prepare_stream(*c, st));
  EDLGT_HIP(hipStreamSynchronize(st));  // the host arrays may go once this returns
  *out = c.release();
  return EDLGT_OK;
}

int edlgt_csr_momentum(const edlgt_csr* h, const edlgt_momentum* m, void* stream, edlgt_csr** out) {
  EDLGT_REQUIRE(h && m && out, "null argument");
  EDLGT_REQUIRE(!h->mf, "not available for a matrix-free operator (no stored entries)");
  *out = nullptr;
  auto c = std::make_unique<edlgt_csr>();
  EDLGT_TRY(csr_momentum(*h, *m, (hipStream_t)stream, c.get()));
  *out = c.release();
  return EDLGT_OK;
}

int edlgt_momentum_destroy(edlgt_momentum* m) {
  delete m;
  return EDLGT_OK;
}

struct edlgt_krylov {
  KrylovSession* s = nullptr;
  const edlgt_basis* b = nullptr;
  int64_t n = 0;
  ~edlgt_krylov() { krylov_destroy(s); }
};

int edlgt_krylov_create(edlgt_csr* h, const edlgt_basis* b, int32_t precision, int32_t m, int32_t reorth,
                        int32_t d_max, void* stream, edlgt_krylov** out) {
  EDLGT_REQUIRE(h && out, "null argument");
  EDLGT_REQUIRE(m >= 1, "m must be >= 1");
  EDLGT_REQUIRE(precision == EDLGT_PREC_FP64 || precision == EDLGT_PREC_MIXED, "bad precision");
  EDLGT_REQUIRE(!b || b->n_states == h->n_rows, "basis and matrix sizes differ");
  EDLGT_REQUIRE(!b || d_max < 1 || b->n_sites * d_max <= 4096, "n_sites * d_max must be <= 4096");
  *out = nullptr;
  auto k = std::make_unique<edlgt_krylov>();
  EDLGT_TRY(krylov_create(*h, b, precision, m, reorth, b ? d_max : 0, (hipStream_t)stream, &k->s));
  k->b = b;
  k->n = h->n_rows;
  *out = k.release();
  return EDLGT_OK;
}

int edlgt_krylov_run(edlgt_krylov* k, void* psi, double dt, int64_t n_steps, double* out_marginals,
                     double* out_err, void* psi_history) {
  EDLGT_REQUIRE(k && psi, "null argument");
  EDLGT_REQUIRE(n_steps >= 0, "n_steps must be >= 0");
  return krylov_session_run(*k->s, psi, dt, n_steps, out_marginals, out_err, psi_history);
}

int edlgt_krylov_destroy(edlgt_krylov* k) {
  delete k;
  return EDLGT_OK;
}

struct edlgt_server {
  void* impl = nullptr;
  int32_t n_eigs = 0;
  int32_t max_slots = 0;
  int32_t n_terms = 1;
};

int edlgt_server_create(edlgt_csr* h, int32_t vec_type, int32_t max_slots, int32_t n_eigs, double tol,
                        int32_t max_basis, int64_t max_matvecs, int32_t warm_cache, void* stream,
                        edlgt_server** out) {
  EDLGT_REQUIRE(h && out, "null argument");
  EDLGT_REQUIRE(!h->mf, "not available for a matrix-free operator (no stored entries)");
  EDLGT_REQUIRE(max_slots >= 1 && max_slots <= 64, "max_slots must be in [1, 64]");
  EDLGT_REQUIRE(n_eigs >= 1 && n_eigs <= h->n_rows && tol > 0.0, "bad n_eigs / tol");
  EDLGT_REQUIRE(vec_type == EDLGT_VALUE_F64 || vec_type == EDLGT_VALUE_C128, "vec_type must be F64 or C128");
  EDLGT_REQUIRE(!(vec_type == EDLGT_VALUE_F64 && h->value_type != EDLGT_VALUE_F64),
                "real vectors need a real matrix");
  EDLGT_REQUIRE(warm_cache >= 0 && warm_cache <= 4096, "warm_cache must be in [0, 4096]");
  *out = nullptr;
  int status = EDLGT_OK;
  void* impl = server_create(*h, vec_type, max_slots, n_eigs, tol, max_basis, max_matvecs, warm_cache,
                             (hipStream_t)stream, &status);
  if (!impl) return status;
  auto* sv = new edlgt_server();
  sv->impl = impl;
  sv->n_eigs = n_eigs;
  sv->max_slots = max_slots;
  sv->n_terms = h->n_terms;
  *out = sv;
  return EDLGT_OK;
}

int edlgt_server_submit(edlgt_server* s, int64_t request_id, const double* coef, uint64_t seed,
                        int32_t* slot_out) {
  EDLGT_REQUIRE(s && coef && slot_out, "null argument");
  int slot = -1;
  EDLGT_TRY(server_submit(s->impl, request_id, coef, seed, &slot));
  *slot_out = slot;
  return EDLGT_OK;
}

int edlgt_server_info(const edlgt_server* s, int32_t* active, int32_t* allocated, int32_t* max_slots,
                      int64_t* slot_bytes, int64_t* shared_bytes) {
  EDLGT_REQUIRE(s, "null server");
  int a = 0, al = 0;
  int64_t sb = 0, wb = 0;
  server_info(s->impl, &a, &al, &sb, &wb);
  if (active) *active = a;
  if (allocated) *allocated = al;
  if (max_slots) *max_slots = s->max_slots;
  if (slot_bytes) *slot_bytes = sb;
  if (shared_bytes) *shared_bytes = wb;
  return EDLGT_OK;
}

int edlgt_server_step(edlgt_server* s, int32_t max_done, int64_t* done_ids, double* done_evals,
                      double* done_resid_est, int64_t* done_matvecs, int32_t* done_converged,
                      int32_t* n_done) {
  EDLGT_REQUIRE(s && done_ids && done_evals && n_done, "null argument");
  int nd = 0;
  const int st = server_step(s->impl, max_done, done_ids, done_evals, done_resid_est, done_matvecs,
                             done_converged, &nd);
  *n_done = nd;
  return st;
}

int edlgt_server_trim(edlgt_server* s) {
  EDLGT_REQUIRE(s, "null server");
  server_trim(s->impl);
  return EDLGT_OK;
}

int edlgt_server_destroy(edlgt_server* s) {
  if (s) {
    server_destroy(s->impl);
    delete s;
  }
  return EDLGT_OK;
}

}  // extern "C"
