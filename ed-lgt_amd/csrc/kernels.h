// Internal entry points shared between the translation units of libedlgt_hip.
#pragma once
#include "common.h"

namespace edlgt {

int sector_build(int32_t n_sites, const int32_t* loc_dims, int32_t d_max, int32_t n_glob,
                 int32_t glob_type, const double* glob_diags, const double* glob_sectors,
                 int32_t n_dirs, const double* link_diags, const double* link_sectors,
                 int32_t n_pairs, const int32_t* pair_dir, const int32_t* pair_site0,
                 const int32_t* pair_site1, int32_t n_nbody, int32_t nbody_type,
                 const double* nbody_diags, const double* nbody_sectors,
                 const int32_t* nb_site_off, const int32_t* nb_sites, hipStream_t st, Basis* out);

int basis_lookup(const Basis& b, const uint16_t* configs_host, int64_t n_query, int64_t* out_idx);

int hamiltonian_build(const Basis& B, int32_t d_max, int32_t n_tables, const double* tables,
                      int32_t n_place, const int32_t* place_nops, const int32_t* place_sites,
                      const int32_t* place_tables, const double* place_strength,
                      const int32_t* place_dagger, const int32_t* place_term, int32_t n_terms,
                      int32_t value_type, hipStream_t st, Csr* out);
// Stored entries of edlgt_hamiltonian_build's CSR from the transition tables
// alone (exact unless two raw entries of a row reach the same ket; never
// below the true count).
int hamiltonian_count(const Basis& B, int32_t d_max, int32_t n_tables, const double* tables, int32_t n_place,
                      const int32_t* place_nops, const int32_t* place_sites, const int32_t* place_tables,
                      const double* place_strength, const int32_t* place_dagger, hipStream_t st,
                      int64_t* out_nnz);
int hamiltonian_build_matfree(const Basis& B, int32_t d_max, int32_t n_tables, const double* tables,
                              int32_t n_place, const int32_t* place_nops, const int32_t* place_sites,
                              const int32_t* place_tables, const double* place_strength,
                              const int32_t* place_dagger, int32_t value_type, hipStream_t st, Csr* out);
int matfree_spmm(const Csr& h, int xtype, int n_rhs, const void* x, int64_t ldx, void* y, int64_t ldy,
                 hipStream_t st);
int placements_expval(const Basis& B, int32_t d_max, int32_t n_tables, const double* tables, int32_t n_place,
                      const int32_t* place_nops, const int32_t* place_sites, const int32_t* place_tables,
                      const double* place_strength, const void* psi, int32_t vec_type, double* out_host,
                      hipStream_t st);

int lanes_per_row(const Csr& h);
// Y[:, b] = H X[:, b] for b < n_rhs; xtype is the element type of X
// (C128/F64, or C64/F32 with precision == MIXED); Y is FP64 (complex iff the
// matrix or X is complex).
// pack (optional, pack_bytes >= spmm_pack_bytes): workspace for the
// row-interleaved FP64 block product (n_rhs > 1, spmm_packable).
int csr_spmm(const Csr& h, int precision, int xtype, int n_rhs, const void* x, int64_t ldx,
             void* y, int64_t ldy, hipStream_t st, void* pack = nullptr, size_t pack_bytes = 0);
size_t spmm_pack_bytes(const Csr& h, int xtype);
bool spmm_packable(const Csr& h, int xtype);
int csr_make_low_precision(Csr& h, hipStream_t st);
// Every eigenpair of a stored single-term matrix (dense.hip): evals (host,
// ascending), evecs (device, N*N values of the matrix's type, row k =
// eigenvector k); at most kDenseEighMaxN states.
constexpr int64_t kDenseEighMaxN = 32768;
int csr_dense_eigh(const Csr& h, double* evals, void* evecs, hipStream_t st);
// Stabilizer Renyi-2 entropy of a complex128 sector state on the support
// holding >= 1 - delta of its weight (magic.hip); at most kSreMaxSupport
// support states.
constexpr int64_t kSreMaxSupport = 16383;
int basis_sre2(const Basis& B, const void* psi, double delta, double* sre2, int64_t* support, double* discarded,
               int64_t* n_strings, hipStream_t st);
// Fused multi-coefficient SpMM (see spmv.hip): xs/ys device pointer arrays,
// coef device [nb][16]; nb in {1, 2, 4, 8, 16}.
constexpr int kCoefMaxRhs = 16;
int coef_pad(int nb);
size_t coef_pack_bytes(const Csr& h, int vec_type, int nb);
bool coef_fused(const Csr& h, int vec_type);
int csr_spmm_coef(const Csr& h, int vec_type, int nb, const void* xs, void* ys, const void* const* hx,
                  void* const* hy, const double* coef, void* pack, hipStream_t st);
// Row partition of the CSR-stream SpMV (run once per matrix).
int csr_prepare_stream(Csr& h, hipStream_t st);
int csr_prepare_tiles(Csr& h, int tile, hipStream_t st);
int csr_compress(Csr& h, hipStream_t st, int* compressed);
int csr_autotune(Csr& h, int precision, int xtype, int n_rhs, const void* x, void* y,
                 hipStream_t st, int32_t* best_g, int64_t* best_grid, float* best_ms);
int convert_f64_to_f32(const double* in, float* out, int64_t n, hipStream_t st);

int lanczos_run(Csr& h, int n_eigs, double tol, int max_basis, int64_t max_iter, int precision,
                int vec_type, const void* v0, double* evals, void* evecs, double* resid,
                int64_t* matvecs, hipStream_t st);

int krylov_run(Csr& h, const Basis* basis, int precision, void* psi, double dt, int64_t n_steps,
               int m, int reorth, int d_max, double* out_marg, double* out_err, void* history,
               hipStream_t st);
int momentum_build(const Basis& B, int n_dims, const int32_t* lvals, const int32_t* unit_cell,
                   const int32_t* k_vals, hipStream_t st, Momentum* out);
int momentum_build_tc(const Basis& B, int d_loc, const int32_t* c_map, const double* c_sign, int32_t k,
                      hipStream_t st, Momentum* out);
int momentum_copy(const Momentum& m, int64_t* col_ptr, int32_t* row_idx, double* data);
int momentum_apply(const Momentum& m, int forward, const void* in, void* out, hipStream_t st);
int csr_momentum(const Csr& h, const Momentum& m, hipStream_t st, Csr* out);

struct KrylovSession;
int krylov_create(Csr& h, const Basis* basis, int precision, int m, int reorth, int d_max, hipStream_t st,
                  KrylovSession** out);
int krylov_session_run(KrylovSession& ks, void* psi, double dt, int64_t n_steps, double* out_marg,
                       double* out_err, void* history);
void krylov_destroy(KrylovSession* ks);

void* server_create(Csr& h, int vec_type, int max_slots, int n_eigs, double tol, int max_basis,
                    int64_t max_mv, int warm_cap, hipStream_t st, int* status);
void server_destroy(void* p);
int server_submit(void* p, int64_t req, const double* coef, uint64_t seed, int* slot);
void server_trim(void* p);
void server_info(void* p, int* active, int* allocated, int64_t* slot_bytes, int64_t* warm_bytes);
int server_step(void* p, int max_done, int64_t* ids, double* evals, double* est, int64_t* mvs,
                int32_t* conv, int* n_done);

int basis_marginals(const Basis& B, const void* psi, bool cplx, int d_max, double* out_dev,
                    hipStream_t st);
// Same result without allocating: partial sums in a static per-device array.
int basis_marginals_noalloc(const Basis& B, const void* psi, bool cplx, int d_max, double* out_dev,
                            hipStream_t st);
// Index and sign of every state's image under a site permutation with per-site
// local maps (host outputs; -1 / an error when an image leaves the sector).
int basis_signed_image(const Basis& B, const int32_t* site_perm, const int32_t* loc_perm, const double* loc_sign,
                       int d_max, int64_t* out_idx, double* out_sign, hipStream_t st);
// y = i^(dir e) x per basis state, e_r = sum_s tab[s][c_r[s]] mod 4 (real or
// complex x, complex y; y may alias a complex x).
// Ensemble density rho[r] = sum_b |sum_{t in block b} coef[t] E[rows[t]][r]|^2
// over the rows of a complex128 matrix E (n_rows_E x n, row-major, device);
// members, coefficients (complex, host) and block starts (host, n_blocks + 1).
int ensemble_density(const void* E, int64_t n, int64_t n_rows_E, int64_t k, const int64_t* rows_host,
                     const double* coef_host, const int64_t* bstart_host, int64_t n_blocks, double* rho_dev,
                     hipStream_t st);
int basis_quarter_phase(const Basis& B, const int32_t* tab_host, int d_max, int n_vec, int dir, bool cplx_in,
                        const void* x, void* y, hipStream_t st);

// out[i] = i-th uniform [0, 1) double of the PCG64 stream with 128-bit state
// and increment {state_hi, state_lo, inc_hi, inc_lo} (numpy's
// default_rng(...).random(n)); complex output carries it in the real part.
int random_uniform(void* out, int64_t n, bool complex_values, const uint64_t state[4], hipStream_t st);

// Schmidt blocks of a state (schmidt.hip): G_b = M_b M_b^H with psi scattered
// to the packed blocks by pos; eigenvalues of Hermitian blocks by Jacobi.
// blocks: host int64 [n_blocks][4] = (d, c, moff, goff).
int schmidt_gram(const void* psi, bool cplx, int64_t n, const int64_t* pos, int n_blocks, const int64_t* blocks,
                 void* gram, int64_t gram_elems, hipStream_t st);
int hermitian_eigvals(void* gram, int64_t gram_elems, int n_blocks, const int64_t* blocks, double* evals,
                      int64_t n_evals, hipStream_t st);

}  // namespace edlgt
