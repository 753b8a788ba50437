/*
 * libedlgt_hip -- MI355X (gfx950) engine for the exact-diagonalization hot
 * path of edlgt (gcataldi96/ed-lgt 0.2.3): symmetry-sector enumeration,
 * Hamiltonian assembly into canonical CSR, SpMV/SpMM, Lanczos and Krylov
 * exp(-iHt).
 *
 * C ABI: plain pointers and sizes only.  Every function returns an int32
 * status (EDLGT_OK == 0, negative on error); edlgt_last_error() gives the
 * message of the most recent failure on the calling thread.
 *
 * Memory: arrays documented as "host" are read or written synchronously
 * through host memory; arrays documented as "device" are device pointers the
 * caller owns (e.g. from hipMalloc or a torch tensor).  Long-lived device
 * state (sector basis, CSR matrix, solver workspaces) lives in opaque handles
 * allocated by the library and released by the matching *_destroy call.  The
 * library never frees caller memory.  `stream` may be NULL (legacy default
 * stream) or a hipStream_t; all work of one call is ordered on that stream.
 * A handle may be used by one thread at a time.
 */
#ifndef EDLGT_HIP_H
#define EDLGT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the Python layer maps them to ValueError/MemoryError/...) */
#define EDLGT_OK 0
#define EDLGT_ERR_INVALID_ARG -1
#define EDLGT_ERR_KEY_NOT_FOUND -2
#define EDLGT_ERR_INDEX_OVERFLOW -3
#define EDLGT_ERR_OUT_OF_MEMORY -4
#define EDLGT_ERR_HIP -5
#define EDLGT_ERR_NOT_CONVERGED -6
#define EDLGT_ERR_NOT_REAL -7

/* ---- value / vector element types */
#define EDLGT_VALUE_F64 0  /* double                       */
#define EDLGT_VALUE_C128 1 /* complex128 as (re, im) double pairs */
#define EDLGT_VALUE_F32 2  /* float                        */
#define EDLGT_VALUE_C64 3  /* complex64 as (re, im) float pairs   */

/* ---- precision modes of SpMV-driven solvers */
#define EDLGT_PREC_FP64 0  /* matrix values and vectors as built (FP64)                     */
#define EDLGT_PREC_MIXED 1 /* FP32 matrix values and gathered vector, FP64 accumulation,
                              dots, recurrence, reorthogonalisation and final refinement   */

/* ---- symmetry types of global / n-body constraints */
#define EDLGT_SYM_U1 0 /* additive: sum of diagonal values equals the sector value        */
#define EDLGT_SYM_Z2 1 /* multiplicative: product of diagonal values equals the sector    */
#define EDLGT_SYM_MOD(n) (16 + (n)) /* additive modulo n (n >= 2), n-body constraints only    */

typedef struct edlgt_basis edlgt_basis; /* sorted sector basis on the device */
typedef struct edlgt_csr edlgt_csr;     /* CSR Hamiltonian on the device     */
typedef struct edlgt_server edlgt_server; /* continuous-batching eigensolver   */
typedef struct edlgt_krylov edlgt_krylov; /* Krylov stepping state (basis)      */
typedef struct edlgt_momentum edlgt_momentum; /* momentum-sector projector     */

int edlgt_last_error(char* buf, int64_t size);
int edlgt_version(void);
/* Device of the calling thread and its free / total memory in bytes. */
int edlgt_device_info(int32_t* device, int64_t* free_bytes, int64_t* total_bytes);

/* ======================================================================
 * Symmetry-sector enumeration.
 *
 * Replaces symmetry_sector_configs (edlgt/symmetries/symmetry_sector.py:68-145,
 * kernel iterative_sitebased_sym_sector_configs :785-967) and
 * get_link_sector_configs (:1167-1234, kernels iterative_link_sector_configs
 * :1237-1326 and iterative_link_sector_configs_plus :1329-1469).
 *
 * Produces every configuration (c_0..c_{n-1}), 0 <= c_s < loc_dims[s], that
 * satisfies all constraints, in lexicographic order (site 0 most
 * significant) -- the order of the reference's sector_configs rows:
 *   global:  sum_s glob_diags[g][s][c_s] == glob_sectors[g]   (EDLGT_SYM_U1)
 *            prod_s glob_diags[g][s][c_s] == glob_sectors[g]  (EDLGT_SYM_Z2)
 *   link:    link_diags[d][0][i][c_i] + link_diags[d][1][j][c_j] == link_sectors[d]
 *            for every pair p (d = pair_dir[p], i = pair_site0[p], j = pair_site1[p])
 *   n-body:  sum (or product) of nbody_diags[k][s][c_s] over the sites
 *            nbody_sites[nbody_site_offsets[k] .. nbody_site_offsets[k+1]) ==
 *            nbody_sectors[k]
 * each within 1e-13 (symmetry_sector.py:17).  All table pointers are host
 * arrays; diag tables are padded to d_max states per site.  n_glob, n_pairs,
 * n_nbody may be 0 (their pointers may then be NULL).  nbody_type may also be
 * EDLGT_SYM_MOD(n): additive modulo n (sum == sector mod n; integer values),
 * the flux constraint of Z_n gauge theories.  Requires loc_dims[s] <= 255 and
 * prod(loc_dims) < 2^120; keys are 128-bit above 2^64 (then the global
 * constraints must be additive with dyadic-rational values).
 * ====================================================================== */
int edlgt_sector_build(int32_t n_sites, const int32_t* loc_dims, int32_t d_max,
                       int32_t n_glob, int32_t glob_type, const double* glob_diags,
                       const double* glob_sectors,
                       int32_t n_dirs, const double* link_diags, const double* link_sectors,
                       int32_t n_pairs, const int32_t* pair_dir, const int32_t* pair_site0,
                       const int32_t* pair_site1,
                       int32_t n_nbody, int32_t nbody_type, const double* nbody_diags,
                       const double* nbody_sectors, const int32_t* nbody_site_offsets,
                       const int32_t* nbody_sites,
                       void* stream, edlgt_basis** out);
int edlgt_basis_info(const edlgt_basis* b, int64_t* n_states, int32_t* n_sites);
/* Sector configurations as uint16 rows (host, n_states x n_sites), the layout of
 * the reference's sector_configs. */
int edlgt_basis_copy_configs(const edlgt_basis* b, uint16_t* out_host);
/* Mixed-radix keys, key = sum_s c_s * prod_{t>s} loc_dims[t], ascending
 * (edlgt/tools/magic.py:40-80 convention).  64-bit when prod(loc_dims) < 2^64,
 * else 128-bit (little-endian word pairs); out_host holds key_words * n_states
 * uint64 words. */
int edlgt_basis_key_words(const edlgt_basis* b, int32_t* key_words);
int edlgt_basis_copy_keys(const edlgt_basis* b, uint64_t* out_host);
/* Index of each query configuration (host uint16 rows) in the basis; replaces
 * config_to_index_binarysearch (edlgt/symmetries/generate_configs.py:145-173).
 * Misses are written as -1 and reported as EDLGT_ERR_KEY_NOT_FOUND. */
int edlgt_basis_lookup(const edlgt_basis* b, const uint16_t* configs_host, int64_t n_query,
                       int64_t* out_idx_host);
/* Device pointers of the keys (uint64 [n]) and configs (uint8 [n][cfg_stride]). */
int edlgt_basis_device_ptrs(const edlgt_basis* b, void** keys, void** configs,
                            int32_t* cfg_stride);
int edlgt_basis_destroy(edlgt_basis* b);

/* ======================================================================
 * Hamiltonian assembly into canonical CSR.
 *
 * Replaces nbody_term and its kernels (edlgt/symmetries/sym_qmb_operations.py:
 * 32-602), the term drivers LocalTerm/TwoBodyTerm/PlaquetteTerm/NBodyTerm
 * .get_Hamiltonian (edlgt/modeling/{local,twobody,plaquette,nbody}_term.py) and QMB_hamiltonian.add_term /
 * build("sparse") (edlgt/modeling/qmb_hamiltonian.py:155-225).
 *
 * tables: host, n_tables dense d_max x d_max complex matrices (interleaved
 *   re, im), each one site-projected local operator O[bra][ket].
 * A placement is one term instance on concrete sites (one site for a local
 *   term, the bond or plaquette sites otherwise):
 *   place_nops[p]          number of factors (1..8); 1 means a diagonal local
 *                          term (localbody_data_par semantics: keep
 *                          |O[a][a]| >= 1e-10 before scaling)
 *   place_sites[8p+f]      site of factor f
 *   place_tables[8p+f]     table of factor f
 *   place_strength[2p..]   complex strength (re, im)
 *   place_dagger[p]        0: emit <bra| s * prod_f O_f |ket>;
 *                          1: emit the Hermitian-conjugate block of the same
 *                          placement (the reference's add_dagger rows)
 * Factor transitions keep |O[a][b]| > 1e-10 (sym_qmb_operations.py:21,85).
 * The result is the canonical CSR of the summed triplets: rows ascending,
 * columns ascending within a row, duplicates summed, explicit zeros kept.
 * value_type EDLGT_VALUE_F64 requires every value's |Im| <= 1e-12
 * (EDLGT_ERR_NOT_REAL otherwise); EDLGT_VALUE_C128 keeps complex values.
 * A ket configuration outside the sector is EDLGT_ERR_KEY_NOT_FOUND.
 * ====================================================================== */
int edlgt_hamiltonian_build(const edlgt_basis* b, int32_t d_max, int32_t n_tables,
                            const double* tables, int32_t n_place, const int32_t* place_nops,
                            const int32_t* place_sites, const int32_t* place_tables,
                            const double* place_strength, const int32_t* place_dagger,
                            int32_t value_type, void* stream, edlgt_csr** out);

/* Number of entries edlgt_hamiltonian_build would store for the same
 * arguments, from the transition tables alone (one thread per row, no
 * enumeration or search): exact unless two raw entries of one row reach the
 * same ket (then an upper bound).  Lets a caller size the CSR, or choose the
 * matrix-free operator, before building.  Replaces the reference's counting
 * pass (_count_row_nnz_from_transition_data + _get_nnz_offsets,
 * edlgt/symmetries/sym_qmb_operations.py:95-164), which counts raw entries
 * before coalescing; this counts the stored ones. */
int edlgt_hamiltonian_count(const edlgt_basis* b, int32_t d_max, int32_t n_tables, const double* tables,
                            int32_t n_place, const int32_t* place_nops, const int32_t* place_sites,
                            const int32_t* place_tables, const double* place_strength,
                            const int32_t* place_dagger, void* stream, int64_t* out_nnz);

/* The same operator without storing it (for sectors whose CSR would not fit
 * in device memory): the handle holds the term program (same arguments as
 * edlgt_hamiltonian_build) and a bucket index over the sorted keys, and every
 * product generates row r's entries on the fly and applies them,
 * y_r = sum over placements and transitions of strength * amp * x[col]
 * (one wave per row, placements over lanes, fixed-order sums).  Products go
 * through edlgt_spmv / edlgt_spmm / edlgt_lanczos / edlgt_krylov_* like a
 * stored matrix (FP64 only; complex128, or float64 for a real operator);
 * entry points that read stored entries (copy, compress, autotune, servers,
 * momentum projection) return EDLGT_ERR_INVALID_ARG.  The first product checks
 * (synchronously) that every ket lies in the sector.  Replaces the
 * ham_format="linear" LinearOperator seam of QMB_hamiltonian.build
 * (edlgt/modeling/qmb_hamiltonian.py:186-216) with an operator that needs
 * no triplets at all. */
int edlgt_hamiltonian_build_matfree(const edlgt_basis* b, int32_t d_max, int32_t n_tables,
                                    const double* tables, int32_t n_place, const int32_t* place_nops,
                                    const int32_t* place_sites, const int32_t* place_tables,
                                    const double* place_strength, const int32_t* place_dagger,
                                    int32_t value_type, void* stream, edlgt_csr** out);
/* *out = 1 for a matrix-free operator, else 0. */
int edlgt_csr_is_matrix_free(const edlgt_csr* h, int32_t* out);
/* Multi-term variant: place_term[p] in [0, n_terms) tags every placement; the
 * result is H = sum_t H_t on one pattern whose entries keep their term id
 * (entries of different terms on the same (row, col) stay separate, sorted by
 * term).  Products then take per-right-hand-side coefficients c[t]
 * (edlgt_spmm_coef, edlgt_server_*), replacing one assembly per parameter
 * value (H(g) = sum_t c_t(g) H_t).  n_terms <= 16.  No reference
 * counterpart: the reference re-assembles H for every coupling. */
int edlgt_hamiltonian_build_terms(const edlgt_basis* b, int32_t d_max, int32_t n_tables,
                                  const double* tables, int32_t n_place, const int32_t* place_nops,
                                  const int32_t* place_sites, const int32_t* place_tables,
                                  const double* place_strength, const int32_t* place_dagger,
                                  const int32_t* place_term, int32_t n_terms, int32_t value_type,
                                  void* stream, edlgt_csr** out);
/* Term ids (host uint8 [nnz], only when n_terms > 1) and the term count. */
int edlgt_csr_copy_terms(const edlgt_csr* h, uint8_t* term_host, int32_t* n_terms);
int edlgt_csr_info(const edlgt_csr* h, int64_t* n_rows, int64_t* nnz, int32_t* value_type);
/* Dictionary coding of the values: when the matrix has at most 4096 distinct
 * (value, term) pairs (compared bitwise), every entry gets a uint8 (<= 256)
 * or uint16 code into a small table, and from then on every FP64 product
 * (edlgt_spmv / edlgt_spmm / edlgt_spmm_coef, the eigensolvers, the servers)
 * streams col + code (5-6 bytes per entry instead of 12-20) and multiplies by
 * exactly the stored values.  *n_dict = 0: too many distinct values, the
 * matrix stays as it was.  FP64 matrices only.  No reference counterpart
 * (the reference stores complex128 triplets / CSC). */
int edlgt_csr_compress(edlgt_csr* h, void* stream, int32_t* n_dict, int32_t* code_bytes);
/* Frees the value (and term-id and reduced-precision) arrays of a coded
 * matrix: products keep working from the codes; edlgt_csr_copy can then no
 * longer return values.  For matrices whose full-width values would crowd
 * out the Krylov vectors (e.g. 1D SU(2), 20 sites: 3.8e8 states). */
int edlgt_csr_release_values(edlgt_csr* h);
/* Copy CSR arrays to host: row_ptr int64 [n+1], col int32 [nnz], val [nnz]. */
int edlgt_csr_copy(const edlgt_csr* h, int64_t* row_ptr, int32_t* col, void* val);
int edlgt_csr_device_ptrs(const edlgt_csr* h, void** row_ptr, void** col, void** val);
/* Build an FP32 (complex64 / float32) copy of the values for EDLGT_PREC_MIXED. */
int edlgt_csr_make_low_precision(edlgt_csr* h, void* stream);
/* Every eigenpair of a stored single-term matrix of up to 32768 states, the
 * dense route of QMB_hamiltonian.diagonalize(n_eigs="full")
 * (edlgt/modeling/qmb_hamiltonian.py:293-301: the dense array factored by
 * scipy.linalg.eigh), here on the device (rocSOLVER zheevd / dsyevd).
 * evals: host double [n] ascending; evecs: device [n * n] of the matrix's
 * value type (complex128 or float64), row k = eigenvector k.  A
 * matrix-free, value-released or multi-term matrix is EDLGT_ERR_INVALID_ARG. */
int edlgt_csr_dense_eigh(const edlgt_csr* h, double* evals, void* evecs, void* stream);
/* A matrix from host CSR arrays (n rows and columns, row_ptr int64 [n+1]
 * from 0 to nnz, col int32 [nnz] in [0, n), val complex128 or float64 [nnz]):
 * the device form of QMB_hamiltonian.add_term + build("sparse")
 * (edlgt/modeling/qmb_hamiltonian.py:155-225) for terms a caller assembles
 * itself; the caller passes SciPy's coalesced form (duplicates summed).
 * Arrays that are not such a CSR are EDLGT_ERR_INVALID_ARG. */
int edlgt_csr_create(int64_t n, int64_t nnz, const int64_t* row_ptr, const int32_t* col, const void* val,
                     int32_t value_type, void* stream, edlgt_csr** out);
int edlgt_csr_destroy(edlgt_csr* h);

/* ======================================================================
 * Sparse products.  Replace csr_spmv / csr_matmat (edlgt/modeling/
 * qmb_hamiltonian.py:705-764) behind the LinearOperator matvec/matmat seam
 * (:186-216).  x, X are device arrays of vec_type: EDLGT_VALUE_C128 or
 * EDLGT_VALUE_F64 (F64 only for real matrices); with EDLGT_PREC_MIXED also
 * EDLGT_VALUE_C64 / EDLGT_VALUE_F32 (FP32 gathered vector).  y, Y are FP64
 * (complex128 for a complex matrix or complex vectors, else float64).  X, Y
 * are n_rhs vectors of length n stored one after the other (column-major
 * N x n_rhs).  precision EDLGT_PREC_MIXED reads the FP32 value copy (built on
 * demand); accumulation is FP64 in every mode.
 * ====================================================================== */
int edlgt_spmv(edlgt_csr* h, int32_t precision, const void* x, void* y, int32_t vec_type,
               void* stream);
int edlgt_spmm(edlgt_csr* h, int32_t precision, int32_t n_rhs, const void* X, void* Y,
               int32_t vec_type, void* stream);

/* Y[:, j] = (sum_t coef_host[j * n_terms + t] H_t) X[:, j] for j < n_rhs on a
 * multi-term matrix (edlgt_hamiltonian_build_terms; a single-term matrix has
 * n_terms = 1).  X, Y: device, column-major with leading dimension n_rows,
 * complex128 (C128) or, for a real matrix, float64 (F64).  Every column's
 * result is independent of n_rhs and of the other columns, bit for bit.
 * Synchronous with respect to stream.  No reference counterpart: it fuses the
 * reference's one-matrix-per-coupling products of a parameter sweep. */
int edlgt_spmm_coef(edlgt_csr* h, int32_t n_rhs, const void* X, void* Y, int32_t vec_type,
                    const double* coef_host, void* stream);

/* Times the SpMM launch shapes with HIP events and keeps the fastest for this
 * matrix (no counterpart in the reference: CPU csr_spmv has no launch shape):
 * for one right-hand side the stream kernel's variants (threads per block x
 * nonzeros per tile x gathers in flight) against the vector kernel's lanes
 * per row x workgroups.  Outputs may be NULL.  A stream winner is reported as
 * *out_lanes = 0, *out_grid = -(variant + 1); a coded matrix
 * (edlgt_csr_compress) only has stream candidates. */
int edlgt_csr_autotune(edlgt_csr* h, int32_t precision, int32_t vec_type, int32_t n_rhs,
                       const void* X, void* Y, void* stream, int32_t* out_lanes, int64_t* out_grid,
                       float* out_ms);
/* Sets the launch shape explicitly (lanes_per_row 0 and grid 0 restore the
 * defaults; lanes_per_row 0 and grid -(v + 1) select stream variant v,
 * 0 <= v < 5, of the coded kernel on a coded matrix, else of the plain one). */
int edlgt_csr_set_launch(edlgt_csr* h, int32_t lanes_per_row, int64_t grid);

/* ======================================================================
 * Lowest eigenpairs by thick-restart Lanczos with full reorthogonalisation.
 * Replaces scipy eigsh(H, k, which="SA") in QMB_hamiltonian.diagonalize
 * (edlgt/modeling/qmb_hamiltonian.py:260-332).
 *   v0           start vector (vec_type, length n) in host or device memory;
 *                NULL -> deterministic pseudo-random entries
 *   tol          stop when every wanted residual ||H y - theta y|| <= tol * max(1,|theta|)
 *   max_basis    Krylov basis size m (>= n_eigs + 2); <= 0 selects
 *                max(2 n_eigs + 1, 20), eigsh's default ncv, and at least
 *                16 blocks of min(n_eigs, 4) vectors when n_eigs > 1
 *   max_iter     total matvec budget
 *   precision    EDLGT_PREC_FP64, or EDLGT_PREC_MIXED (FP32 SpMV, then FP64
 *                refinement restarted from the FP32 Ritz vectors)
 *   out_evals    host double [n_eigs], ascending total energies
 *   out_evecs    device, vec_type, n x n_eigs column-major, or NULL
 *   out_resid    host double [n_eigs] final FP64 residual norms, or NULL
 *   out_matvecs  host int64 [2]: FP64 and FP32 matvec counts, or NULL
 * Returns EDLGT_ERR_NOT_CONVERGED (with the best estimates written) when the
 * budget is exhausted.
 * ====================================================================== */
int edlgt_lanczos(edlgt_csr* h, int32_t n_eigs, double tol, int32_t max_basis,
                  int64_t max_iter, int32_t precision, int32_t vec_type, const void* v0,
                  double* out_evals, void* out_evecs, double* out_resid,
                  int64_t* out_matvecs, void* stream);

/* Uniform [0, 1) doubles on the device from a PCG64 stream: out[i] is the
 * i-th value of numpy.random.Generator(PCG64).random(n) for the bit
 * generator whose 128-bit state and increment are
 * state = {state_hi, state_lo, inc_hi, inc_lo} (numpy's
 * bit_generator.state["state"]), bit for bit.  value_type EDLGT_VALUE_F64
 * writes doubles, EDLGT_VALUE_C128 writes (value, 0) pairs.
 * No counterpart in the reference's code: its eigsh draws its own random
 * start vector; the model layer here starts every solve from
 * default_rng(0).random(n) (a fixed seed, so solves are reproducible), and
 * this builds it on the device instead of on the host plus a full-length
 * host-to-device copy. */
int edlgt_random_uniform(void* out, int64_t n, int32_t value_type, const uint64_t* state, void* stream);

/* ======================================================================
 * Served lowest-eigenvalue solves with continuous batching (no reference
 * counterpart; each request replaces one build_Hamiltonian +
 * diagonalize_Hamiltonian of the reference at its own couplings).
 * A server owns max_slots solver slots on a multi-term matrix h; a request is
 * a coefficient vector c (n_terms doubles: H(c) = sum_t c[t] H_t) and a seed
 * for its random start block.  edlgt_server_submit puts it in a free slot
 * (*slot_out = -1 when all slots are busy: retry after a step).
 * edlgt_server_step advances every active slot by one block-Lanczos expansion
 * (one fused SpMM streams the matrix once for all slots) and reports the
 * requests that finished: ids, n_eigs ascending eigenvalues each (total
 * energies), residual estimates, matvec counts, converged flags.
 * Slots are allocated on first use, slot_bytes each (edlgt_server_info), so
 * a caller admits by device-memory footprint: submit only while a free
 * allocated slot exists or the budget covers one more slot.
 * warm_cache = 0: a request starts from its seed alone and its result is
 * bitwise independent of the other requests.  warm_cache = C > 0: the server
 * keeps the Ritz vectors of up to C finished requests and starts each new
 * request from those of the nearest coefficient vector; results then agree
 * with the cold ones to tol.  shared_bytes = that cache plus the packed
 * right-hand-side workspace of the fused SpMM (both allocated on first use).
 * ====================================================================== */
int edlgt_server_create(edlgt_csr* h, int32_t vec_type, int32_t max_slots, int32_t n_eigs, double tol,
                        int32_t max_basis, int64_t max_matvecs, int32_t warm_cache, void* stream,
                        edlgt_server** out);
int edlgt_server_submit(edlgt_server* s, int64_t request_id, const double* coef, uint64_t seed,
                        int32_t* slot_out);
int edlgt_server_info(const edlgt_server* s, int32_t* active, int32_t* allocated, int32_t* max_slots,
                      int64_t* slot_bytes, int64_t* shared_bytes);
int edlgt_server_step(edlgt_server* s, int32_t max_done, int64_t* done_ids, double* done_evals,
                      double* done_resid_est, int64_t* done_matvecs, int32_t* done_converged,
                      int32_t* n_done);
/* Frees the device memory of every idle slot (for a budget shared with other
 * servers); busy slots and the warm cache are kept. */
int edlgt_server_trim(edlgt_server* s);
int edlgt_server_destroy(edlgt_server* s);

/* ======================================================================
 * Diagonal local observables.  Per-site marginals of |psi|^2:
 *   out_host[s * d_max + a] = sum_r |psi_r|^2 [config_r[s] == a]
 * so that <psi|O_s|psi> = sum_a out[s][a] O_s[a][a] for every diagonal local
 * operator -- replaces LocalTerm.get_expval + exp_val_data
 * (edlgt/modeling/local_term.py:119-237, edlgt/modeling/qmb_state.py:630-653)
 * for diagonal operators.  psi: device, vec_type C128 or F64, ||psi|| <= 1.
 * Deterministic (fixed-point integer accumulation, |error| <= N * 2^-62).
 * ====================================================================== */
int edlgt_basis_marginals(const edlgt_basis* b, const void* psi, int32_t vec_type, int32_t d_max,
                          double* out_host, void* stream);
/* Allocation-free variant of edlgt_basis_marginals.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * out_dev: device double [n_sites * d_max], written in stream order (no
 * synchronisation, no host copy).  The per-block partial sums live in a
 * static 4 MiB array per device (the library's only fixed scratch; held until
 * the process ends), so n_sites * d_max <= 4096 as above and nothing else is
 * limited.  The result equals edlgt_basis_marginals' bit for bit.  Calls that
 * overlap on one device share the array and race: order them on one stream. */
int edlgt_basis_marginals_noalloc(const edlgt_basis* b, const void* psi, int32_t vec_type, int32_t d_max,
                                  double* out_dev, void* stream);

/* Ensemble density of eigenvector rows: density[r] (device, n) =
 * sum_b |sum_{t in block b} member_coef[t] * rows_matrix[member_rows[t]][r]|^2
 * for a complex128 row-major matrix (device, n_rows x n); members,
 * coefficients (complex128 as re, im pairs) and block_start (n_blocks + 1,
 * from 0 to n_members) on the host.  The weighted sums behind
 * microcanonical_avg / diagonal_avg / canonical_avg
 * (edlgt/models/quantum_model.py:912-1156): equal weights over the shell, the
 * state's projection on each degenerate level, Boltzmann weights (one member
 * per block, coefficient sqrt(w)).  Rows outside the matrix or block starts
 * that are not a partition are EDLGT_ERR_INVALID_ARG. */
int edlgt_ensemble_density(const void* rows_matrix, int64_t n, int64_t n_rows, int64_t n_members,
                           const int64_t* member_rows, const double* member_coef, const int64_t* block_start,
                           int64_t n_blocks, double* density, void* stream);

/* Stabilizer Renyi-2 entropy of a complex128 state psi (device, n_states) on
 * the support holding >= 1 - prob_threshold of its weight: replaces
 * QMB_state.stabilizer_renyi_entropy (edlgt/modeling/qmb_state.py:393-465,
 * with extract_support :708-790 and edlgt/tools/magic.py:283-649).  Outputs
 * (host): the SRE2, the support size, the discarded weight and the number of
 * X-strings.  prob_threshold outside (0, 1), a support above 16383 states or
 * keys above 64 bits are EDLGT_ERR_INVALID_ARG. */
int edlgt_basis_sre2(const edlgt_basis* b, const void* psi, double prob_threshold, double* sre2,
                     int64_t* support_size, double* discarded_weight, int64_t* n_strings, void* stream);

/* Signed image of every basis state under a site permutation with per-site
 * local maps: the image of configuration r puts loc_perm[j * d_max + c_r[j]]
 * on site site_perm[j], with sign prod_j loc_sign[j * d_max + c_r[j]];
 * out_index[r] = its basis index, out_sign[r] the sign (host arrays of N).
 * Replaces build_parity_operator / parity_image_config
 * (edlgt/symmetries/inversion_sym.py:77-161) behind
 * SU2_Model.get_parity_inversion_operator (edlgt/models/SU2_model.py:708-763).
 * An image outside the sector is EDLGT_ERR_KEY_NOT_FOUND (the reference
 * writes -1 there); site_perm must be a permutation and loc_perm[j] must map
 * into site site_perm[j]'s local basis (EDLGT_ERR_INVALID_ARG otherwise). */
int edlgt_basis_signed_image(const edlgt_basis* b, const int32_t* site_perm, const int32_t* loc_perm,
                             const double* loc_sign, int32_t d_max, int64_t* out_index, double* out_sign,
                             void* stream);

/* Diagonal quarter-turn phase of the basis states:
 *   y[v][r] = i^(e_r) x[v][r],  e_r = (sum_s site_exp[s * d_max + config_r[s]]) mod 4
 * x: device float64 (n_vec x N), y: device complex128 (n_vec x N).  The
 * real-gauge solve of QED with staggered matter (QED_Model real_gauge=True)
 * diagonalises D^H H D, D = diag(i^e), which is real; this maps its real
 * eigenvectors back to eigenvectors of H.  No counterpart in the reference
 * (its QED-with-matter solve, eigsh in QMB_hamiltonian.diagonalize,
 * edlgt/modeling/qmb_hamiltonian.py:260-332, stays complex); the results are
 * the same eigenpairs.  site_exp: host int32 [n_sites * d_max]. */
int edlgt_basis_quarter_phase(const edlgt_basis* b, const int32_t* site_exp, int32_t d_max, int32_t n_vec,
                              const double* x, void* y, void* stream);
/* The same phase on complex vectors, either way: y[v][r] = i^(direction e_r)
 * x[v][r], direction +1 (D x) or -1 (D^H x); x, y: device complex128
 * (n_vec x N), y may equal x.  The real-gauge time evolution
 * (time_evolution_Hamiltonian(..., real_gauge=True)) evolves D^H psi0 under
 * the real D^H H D and maps the states back; the reference's expm_multiply on
 * the complex matrix (edlgt/modeling/qmb_hamiltonian.py:334-426) gives the
 * same states.  No counterpart in the reference. */
int edlgt_basis_quarter_phase_c(const edlgt_basis* b, const int32_t* site_exp, int32_t d_max, int32_t n_vec,
                                int32_t direction, const void* x, void* y, void* stream);

/* ======================================================================
 * Schmidt blocks of a state over a bipartition of the lattice.  Replaces the
 * psi_matrix reshape of QMB_state._get_psi_matrix and the dense products and
 * SVD of reduced_density_matrix / entanglement_entropy
 * (edlgt/modeling/qmb_state.py:160-291; partition maps of
 * QuantumModel._get_partition, edlgt/models/quantum_model.py:715-791).
 * In a symmetry sector psi[subsystem config, environment config] is block
 * diagonal up to permutations (connected components of the pairs present);
 * the caller lays the blocks out: blocks (host int64 [n_blocks][4]) =
 * (d, c, moff, goff), block b a dense d x c row-major matrix M_b at element
 * offset moff of a packed array, and state i at element pos[i] (pos: device
 * int64 [n_states]; a position outside the blocks is EDLGT_ERR_INVALID_ARG).
 *
 * edlgt_schmidt_gram: gram (device complex128, gram_elems elements) receives
 * G_b = M_b M_b^H at offset goff (d x d row-major, exactly Hermitian).  psi:
 * device, vec_type C128 or F64.  With rows = the subsystem side, G_b is the
 * block of the reduced density matrix; with rows = the smaller side its
 * eigenvalues are the squared Schmidt values.
 *
 * edlgt_hermitian_eigvals: eigenvalues of the Hermitian blocks of gram (same
 * block table, only d and goff read) by cyclic Jacobi on the device;
 * evals (host double, n_evals >= sum of d) receives block b's d values at
 * offset d_0 + ... + d_{b-1}, unsorted.  gram is overwritten.
 * EDLGT_ERR_NOT_CONVERGED after 60 sweeps.
 * ====================================================================== */
int edlgt_schmidt_gram(const void* psi, int32_t vec_type, int64_t n_states, const int64_t* pos, int32_t n_blocks,
                       const int64_t* blocks, void* gram, int64_t gram_elems, void* stream);
int edlgt_hermitian_eigvals(void* gram, int64_t gram_elems, int32_t n_blocks, const int64_t* blocks, double* evals,
                            int64_t n_evals, void* stream);

/* ======================================================================
 * Expectation values of term instances without building their matrices:
 *   out_host[2p], out_host[2p+1] = Re, Im of <psi| strength_p prod_f O_f |psi>
 * for every placement p (same tables / placement arrays as
 * edlgt_hamiltonian_build, no dagger flags: each placement is one operator
 * product on its sites, bra transitions O[bra][ket], transitions below 1e-10
 * dropped, single-factor placements diagonal-only as in the build).
 * Replaces the per-instance nbody_term + exp_val_data passes of
 * lattice_twobody_exp_val (edlgt/modeling/twobody_term.py:232-264),
 * PlaquetteTerm.get_expval (edlgt/modeling/plaquette_term.py:165-263) and
 * NBodyTerm.get_expval (edlgt/modeling/nbody_term.py:164-220) with
 * exp_val_data (edlgt/modeling/qmb_state.py:630-653): one pass over the
 * basis serves every placement.  psi: device, vec_type C128 or F64.  A ket
 * outside the sector is an error (EDLGT_ERR_KEY_NOT_FOUND).  Deterministic
 * (fixed summation order).
 * ====================================================================== */
int edlgt_expval_terms(const edlgt_basis* b, int32_t d_max, int32_t n_tables, const double* tables,
                       int32_t n_place, const int32_t* place_nops, const int32_t* place_sites,
                       const int32_t* place_tables, const double* place_strength, const void* psi,
                       int32_t vec_type, double* out_host, void* stream);

/* ======================================================================
 * Momentum sectors.  edlgt_momentum_build forms the projector B (n_rows x
 * n_cols) onto momentum k of the block translations by unit_cell[d] sites
 * along each axis of lvals (n_dims axes; axis order and site coordinates as
 * the reference's translational_sym module: row-major, last axis fastest):
 * one column per translation orbit compatible with k (its smallest basis
 * index, ascending), the normalised Bloch sum over the orbit's period box,
 * equal images accumulated, |amplitude| <= 1e-14 dropped for k != 0 (a
 * cancelled column stays, empty).  Replaces get_momentum_basis
 * (edlgt/symmetries/translational_sym.py:1307-1359, :568-950) used by
 * QuantumModel.set_momentum_sector (edlgt/models/quantum_model.py:213-276);
 * TC_symmetry is not supported.  The sector must be closed under the
 * translations (else EDLGT_ERR_KEY_NOT_FOUND).
 *   edlgt_momentum_copy: CSC of B to the host (col_ptr [n_cols+1], row_idx
 *                        [nnz], data [nnz] complex as (re, im) pairs)
 *   edlgt_momentum_apply: forward 1: out (n_rows) = B in (n_cols); 0: out
 *                        (n_cols) = B^H in (n_rows); complex128 device vectors
 *   edlgt_csr_momentum:  new CSR B^H H B (n_cols x n_cols, canonical: columns
 *                        sorted, duplicates summed), float64 when H is real
 *                        and k = 0 -- replaces the momentum triplet builders
 *                        nbody_data_momentum* (translational_sym.py:1394-1969)
 *                        behind QMB_hamiltonian.build; single-term, values kept
 * ====================================================================== */
int edlgt_momentum_build(const edlgt_basis* b, int32_t n_dims, const int32_t* lvals, const int32_t* unit_cell,
                         const int32_t* k_vals, void* stream, edlgt_momentum** out);
/* The same for the translation-combined symmetry of a 1D ring: X = shift by
 * one site after the local involution c_map (d_loc entries), momentum k of
 * the L-fold cyclic group of X; an odd power of X carries the reference
 * configuration's sign prod_j c_sign[config_j]; the Bloch phases are
 * exp(-2 pi i k t / L); at k = 0 a column whose norm^2 is within 1e-8 of zero
 * stays empty.  Replaces build_TC_translations / momentum_basis_zero_k_TC /
 * momentum_basis_finite_k_TC (translational_sym.py:111-128, 457-505,
 * 953-1306), which hold c_map and c_sign for the SU(2) chain. */
int edlgt_momentum_build_tc(const edlgt_basis* b, int32_t d_loc, const int32_t* c_map, const double* c_sign,
                            int32_t k, void* stream, edlgt_momentum** out);
int edlgt_momentum_info(const edlgt_momentum* m, int64_t* n_rows, int64_t* n_cols, int64_t* nnz,
                        int32_t* is_complex);
int edlgt_momentum_copy(const edlgt_momentum* m, int64_t* col_ptr, int32_t* row_idx, double* data);
int edlgt_momentum_apply(const edlgt_momentum* m, int32_t forward, const void* in, void* out, void* stream);
int edlgt_csr_momentum(const edlgt_csr* h, const edlgt_momentum* m, void* stream, edlgt_csr** out);
int edlgt_momentum_destroy(edlgt_momentum* m);

/* ======================================================================
 * Real-time evolution psi <- exp(-i H dt)^n_steps psi by Krylov (Lanczos)
 * exponentiation with an m-dimensional basis per step; replaces
 * QMB_hamiltonian.time_evolution -> expm_multiply(-1j*H, psi0, ...)
 * (edlgt/modeling/qmb_hamiltonian.py:334-426).
 *   psi            device complex128 [n], overwritten by the final state
 *   reorth         1: full CGS2 reorthogonalisation of the Krylov basis; 0: three-term
 *   out_marginals  host [(n_steps+1) * n_sites * d_max] per-step marginals
 *                  (step 0 = input state), or NULL (then b may be NULL)
 *   out_err        host [n_steps] a-posteriori Krylov error estimates, or NULL
 *   psi_history    device complex128 [(n_steps+1) * n] states per step, or NULL
 * precision EDLGT_PREC_MIXED runs the matvecs with FP32 values and an FP32
 * copy of the Krylov vector; the recurrence and exponential stay FP64.
 * ====================================================================== */
int edlgt_krylov_evolve(edlgt_csr* h, const edlgt_basis* b, int32_t precision, void* psi, double dt,
                        int64_t n_steps, int32_t m, int32_t reorth, int32_t d_max,
                        double* out_marginals, double* out_err, void* psi_history, void* stream);

/* The same stepping with its state kept between calls: create allocates the
 * m+1 basis vectors (and the marginal buffer when b is given and d_max >= 1)
 * once; each run advances psi by n_steps steps of dt exactly as
 * edlgt_krylov_evolve does (same arguments, same results); destroy frees
 * them.  Lets a caller evolve a trajectory in chunks (observables, progress,
 * checkpoints between chunks) without re-allocating the basis, which for
 * hundreds of millions of states is a hundred-GB allocation.  The session
 * uses h (and b) until destroyed; one host thread at a time.
 * Replaces the same expm_multiply call as edlgt_krylov_evolve
 * (edlgt/modeling/qmb_hamiltonian.py:334-426). */
int edlgt_krylov_create(edlgt_csr* h, const edlgt_basis* b, int32_t precision, int32_t m, int32_t reorth,
                        int32_t d_max, void* stream, edlgt_krylov** out);
int edlgt_krylov_run(edlgt_krylov* k, void* psi, double dt, int64_t n_steps, double* out_marginals,
                     double* out_err, void* psi_history);
int edlgt_krylov_destroy(edlgt_krylov* k);

#ifdef __cplusplus
}
#endif
#endif
