"""Python handles around the device objects of libedlgt_hip.

SectorBasis  -- sorted symmetry-sector basis (keys + configurations) on the GPU
CSRMatrix    -- canonical CSR Hamiltonian on the GPU, with SpMV/SpMM and Lanczos;
                optionally multi-term (H = sum_t H_t, coefficients per product)
Server       -- continuous-batching lowest-eigenvalue solver on a multi-term matrix
"""
import ctypes as C
import weakref

import numpy as np

from . import _lib as L


# id of a configs() table -> its basis, so that API calls given the model's
# sector_configs (as the reference's are) can reach the device basis
_CONFIG_OWNER = {}


def basis_of_configs(configs):
    """The SectorBasis whose configs() returned this very array, or None."""
    ref = _CONFIG_OWNER.get(id(configs))
    b = ref() if ref is not None else None
    return b if b is not None and b._configs is configs else None


def ensemble_density(E, members, coef, block_start, stream=None):
    """rho (device float64, n) = sum_b |sum_{t in block b} coef[t] E[members[t]]|^2
    over the rows of a complex128 device matrix E (edlgt_ensemble_density)."""
    torch = _torch()
    E = E.to(device="cuda", dtype=torch.complex128).contiguous()
    n_rows, n = E.shape
    members = np.ascontiguousarray(members, dtype=np.int64)
    coef = np.ascontiguousarray(coef, dtype=np.complex128)
    bs = np.ascontiguousarray(block_start, dtype=np.int64)
    rho = torch.empty(n, dtype=torch.float64, device="cuda")
    L.check(L.load().edlgt_ensemble_density(L.ptr(E), n, n_rows, members.size, members.ctypes.data,
                                            coef.ctypes.data, bs.ctypes.data, bs.size - 1, L.ptr(rho),
                                            L.stream_handle(stream)))
    return rho


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("edlgt_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


class SectorBasis:
    """Replaces the uint16 `sector_configs` table of the reference
    (edlgt/symmetries/symmetry_sector.py:68-145, 1167-1234) by a device
    object; `configs()` returns the same table on the host."""

    def __init__(self, loc_dims, glob=None, link=None, nbody=None, stream=None):
        """glob = (diags (n_glob, n_sites, d_max), sectors, type 'U'|'Z') or None
        link  = (diags (n_dirs, 2, n_sites, d_max), sectors (n_dirs), pairs [(dir, i, j)]) or None
        nbody = (diags (n_nb, n_sites, d_max), sectors, [site lists], type) or None"""
        lib = L.load()
        self.loc_dims = np.ascontiguousarray(loc_dims, dtype=np.int32)
        n_sites = self.loc_dims.size
        d_max = int(self.loc_dims.max())
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data

        n_glob = glob_type = 0
        gd = gs = None
        if glob is not None:
            g_diags, g_sec, g_type = glob
            g_diags = np.asarray(g_diags, dtype=np.float64)
            d_max = max(d_max, g_diags.shape[-1])
            n_glob = g_diags.shape[0]
            glob_type = L.SYM_U1 if g_type == "U" else L.SYM_Z2
        n_dirs = n_pairs = 0
        ld = ls = pd = p0 = p1 = None
        if link is not None:
            l_diags, l_sec, pairs = link
            l_diags = np.asarray(l_diags, dtype=np.float64)
            d_max = max(d_max, l_diags.shape[-1])
            n_dirs = l_diags.shape[0]
            pairs = np.asarray(pairs, dtype=np.int32).reshape(-1, 3)
            n_pairs = pairs.shape[0]
        n_nb = nb_type = 0
        nd = ns = no = nsi = None
        if nbody is not None:
            nb_diags, nb_sec, nb_sites, nb_t = nbody
            nb_diags = np.asarray(nb_diags, dtype=np.float64)
            d_max = max(d_max, nb_diags.shape[-1])
            n_nb = nb_diags.shape[0]
            nb_type = L.SYM_U1 if nb_t == "U" else L.SYM_Z2 if nb_t == "Z" else int(nb_t)

        def pad(a):
            out = np.zeros(a.shape[:-1] + (d_max,), dtype=np.float64)
            out[..., : a.shape[-1]] = a
            return out

        if glob is not None:
            gd = arr(pad(g_diags), np.float64)
            gs = arr(g_sec, np.float64)
        if link is not None:
            ld = arr(pad(l_diags), np.float64)
            ls = arr(l_sec, np.float64)
            pd = arr(pairs[:, 0], np.int32)
            p0 = arr(pairs[:, 1], np.int32)
            p1 = arr(pairs[:, 2], np.int32)
        if nbody is not None:
            nd = arr(pad(nb_diags), np.float64)
            ns = arr(nb_sec, np.float64)
            offs = np.zeros(n_nb + 1, dtype=np.int32)
            flat = []
            for k, sl in enumerate(nb_sites):
                flat.extend(int(s) for s in sl)
                offs[k + 1] = len(flat)
            no = arr(offs, np.int32)
            nsi = arr(np.array(flat, dtype=np.int32), np.int32)
        h = C.c_void_p()
        st = L.stream_handle(stream) if stream is not None else None
        L.check(lib.edlgt_sector_build(
            n_sites, self.loc_dims.ctypes.data, d_max, n_glob, glob_type, gd, gs, n_dirs, ld, ls,
            n_pairs, pd, p0, p1, n_nb, nb_type, nd, ns, no, nsi, st, C.byref(h)))
        self._h = h
        n = C.c_int64()
        L.check(lib.edlgt_basis_info(h, C.byref(n), None))
        self.n_states = int(n.value)
        self.n_sites = n_sites
        self._configs = None

    @property
    def handle(self):
        return self._h

    def configs(self):
        """uint16 (n_states, n_sites) table, lexicographically sorted."""
        if self._configs is None:
            out = np.empty((self.n_states, self.n_sites), dtype=np.uint16)
            L.check(L.load().edlgt_basis_copy_configs(self._h, out.ctypes.data))
            self._configs = out
            _CONFIG_OWNER[id(out)] = weakref.ref(self)
        return self._configs

    def sre2(self, psi, prob_threshold=1e-2, stream=None):
        """Stabilizer Renyi-2 entropy of a sector state on the support holding
        >= 1 - prob_threshold of its weight (edlgt_basis_sre2): (SRE2,
        support size, discarded weight, number of X-strings)."""
        torch = _torch()
        t = psi if hasattr(psi, "data_ptr") else torch.as_tensor(np.asarray(psi, dtype=np.complex128))
        t = t.to(device="cuda", dtype=torch.complex128).reshape(-1).contiguous()
        if t.numel() != self.n_states:
            raise ValueError(f"the state has {t.numel()} entries, the sector {self.n_states}")
        out, disc = np.zeros(1), np.zeros(1)
        k, ns = np.zeros(1, dtype=np.int64), np.zeros(1, dtype=np.int64)
        L.check(L.load().edlgt_basis_sre2(self._h, L.ptr(t), float(prob_threshold), out.ctypes.data, k.ctypes.data,
                                          disc.ctypes.data, ns.ctypes.data, L.stream_handle(stream)))
        return float(out[0]), int(k[0]), float(disc[0]), int(ns[0])

    def keys(self):
        """Keys as uint64 (n,) or, for 128-bit keys, Python ints."""
        kw = C.c_int32()
        L.check(L.load().edlgt_basis_key_words(self._h, C.byref(kw)))
        out = np.empty(self.n_states * kw.value, dtype=np.uint64)
        L.check(L.load().edlgt_basis_copy_keys(self._h, out.ctypes.data))
        if kw.value == 1:
            return out
        w = out.reshape(-1, 2)
        return [int(lo) | (int(hi) << 64) for lo, hi in w]

    def signed_image(self, site_perm, loc_perm, loc_sign, stream=None):
        """(index, sign) of every state's image under a site permutation with
        per-site local maps: the image puts loc_perm[j][c_j] on site
        site_perm[j] and carries prod_j loc_sign[j][c_j]
        (edlgt_basis_signed_image).  loc_perm / loc_sign: (n_sites, d) or one
        row for every site."""
        sp = np.ascontiguousarray(site_perm, dtype=np.int32)
        lp = np.atleast_2d(np.asarray(loc_perm, dtype=np.int32))
        ls = np.atleast_2d(np.asarray(loc_sign, dtype=np.float64))
        if lp.shape[0] == 1:
            lp = np.repeat(lp, self.n_sites, axis=0)
            ls = np.repeat(ls, self.n_sites, axis=0)
        lp, ls = np.ascontiguousarray(lp), np.ascontiguousarray(ls)
        if sp.shape != (self.n_sites,) or lp.shape != ls.shape or lp.shape[0] != self.n_sites:
            raise ValueError("site_perm needs one entry per site, loc_perm / loc_sign one row per site")
        idx = np.empty(self.n_states, dtype=np.int64)
        sign = np.empty(self.n_states, dtype=np.float64)
        L.check(L.load().edlgt_basis_signed_image(self._h, sp.ctypes.data, lp.ctypes.data, ls.ctypes.data,
                                                  lp.shape[1], idx.ctypes.data, sign.ctypes.data,
                                                  L.stream_handle(stream)))
        return idx, sign

    def quarter_phase(self, site_exp, x, direction=1, out=None, stream=None):
        """y[..., r] = i^(direction e_r) x[..., r] with e_r = sum_s site_exp[s, c_r[s]]
        mod 4, for device vectors x ((N,) or (k, N)): float64 x (direction +1,
        edlgt_basis_quarter_phase) or complex128 x (either direction, out may
        be x itself, edlgt_basis_quarter_phase_c); complex128 result."""
        torch = _torch()
        tab = np.ascontiguousarray(site_exp, dtype=np.int32)
        if tab.ndim != 2 or tab.shape[0] != self.n_sites or tab.shape[1] < int(self.loc_dims.max()):
            raise ValueError("site_exp must be (n_sites, d_max) with d_max >= every local dimension")
        if x.dtype not in (torch.float64, torch.complex128) or not x.is_cuda or x.shape[-1] != self.n_states:
            raise ValueError("x must be a float64 or complex128 CUDA tensor over the sector")
        if not x.is_contiguous():
            raise ValueError("x must be contiguous")
        n_vec = 1 if x.dim() == 1 else int(x.shape[0])
        y = torch.empty(x.shape, dtype=torch.complex128, device=x.device) if out is None else out
        if y.shape != x.shape or y.dtype != torch.complex128 or not y.is_contiguous():
            raise ValueError("out must be a contiguous complex128 tensor shaped like x")
        if x.dtype == torch.float64:
            if direction != 1:
                raise ValueError("a real vector is mapped with direction +1")
            L.check(L.load().edlgt_basis_quarter_phase(self._h, tab.ctypes.data, tab.shape[1], n_vec, L.ptr(x),
                                                       L.ptr(y), L.stream_handle(stream)))
        else:
            L.check(L.load().edlgt_basis_quarter_phase_c(self._h, tab.ctypes.data, tab.shape[1], n_vec,
                                                         int(direction), L.ptr(x), L.ptr(y),
                                                         L.stream_handle(stream)))
        return y

    def lookup(self, configs):
        """Indices of configurations (raises ValueError if one is not in the sector)."""
        q = np.ascontiguousarray(np.atleast_2d(configs), dtype=np.uint16)
        if q.shape[1] != self.n_sites:
            raise ValueError("configuration length does not match the lattice")
        out = np.empty(q.shape[0], dtype=np.int64)
        L.check(L.load().edlgt_basis_lookup(self._h, q.ctypes.data, q.shape[0], out.ctypes.data))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.load().edlgt_basis_destroy(h)
            except Exception:
                pass
            self._h = None


def _pack_placements(placements):
    """Placement dicts {sites, tables, strength[, dagger, term]} -> the C ABI's
    arrays (nops, sites[8], tables[8], strength (re, im), dagger, term)."""
    n_place = len(placements)
    nops = np.zeros(n_place, dtype=np.int32)
    sites = np.zeros((n_place, 8), dtype=np.int32)
    tabs = np.zeros((n_place, 8), dtype=np.int32)
    stren = np.zeros((n_place, 2), dtype=np.float64)
    dag = np.zeros(n_place, dtype=np.int32)
    term = np.zeros(n_place, dtype=np.int32)
    for p, pl in enumerate(placements):
        k = len(pl["sites"])
        if k > 8:
            raise ValueError("a placement acts on at most 8 sites")
        nops[p] = k
        sites[p, :k] = pl["sites"]
        tabs[p, :k] = pl["tables"]
        s = complex(pl.get("strength", 1.0))
        stren[p] = (s.real, s.imag)
        dag[p] = 1 if pl.get("dagger") else 0
        term[p] = int(pl.get("term", 0))
    return nops, sites, tabs, stren, dag, term


def stored_entries_bound(tables, placements, n_states):
    """Host-side upper bound of the stored entries: every placement's largest
    transition count per row (products of the factors' largest row counts of
    entries above 1e-10; single-factor terms are diagonal), times the states."""
    tables = np.asarray(tables)
    row_max = (np.abs(tables) > 1e-10).sum(axis=2).max(axis=1)
    col_max = (np.abs(tables) > 1e-10).sum(axis=1).max(axis=1)
    per_row = 0
    for pl in placements:
        if len(pl["tables"]) == 1:
            per_row += 1
            continue
        m = col_max if pl.get("dagger") else row_max
        per_row += int(np.prod([max(int(m[t]), 0) for t in pl["tables"]]))
    return int(n_states) * per_row


def stored_entries(basis, tables, placements, stream=None):
    """Entries the stored CSR of these placements would hold, counted on the
    device from the transition tables (edlgt_hamiltonian_count: exact unless
    two raw entries of a row reach the same ket, never below the true count)."""
    tables = np.ascontiguousarray(tables, dtype=np.complex128)
    n_tables, d_max = tables.shape[0], tables.shape[1]
    nops, sites, tabs, stren, dag, _ = _pack_placements(placements)
    out = C.c_int64()
    L.check(L.load().edlgt_hamiltonian_count(
        basis.handle, d_max, n_tables, tables.ctypes.data, len(placements), nops.ctypes.data,
        sites.ctypes.data, tabs.ctypes.data, stren.ctypes.data, dag.ctypes.data,
        L.stream_handle(stream) if stream is not None else None, C.byref(out)))
    return int(out.value)


def storage_plan(n_states, nnz, value_type, free_bytes):
    """How to hold a Hamiltonian with nnz stored entries in free_bytes of
    device memory, leaving about half of it to the solvers: "stored" (CSR
    with values), "coded" (CSR built in full, then its values replaced by
    codes of at most 2 bytes), or "free" (matrix-free operator)."""
    s_v = 8 if value_type == L.VALUE_F64 else 16
    full = int(nnz) * (4 + s_v) + 8 * (int(n_states) + 1)
    coded = int(nnz) * (4 + 2) + 8 * (int(n_states) + 1)
    if full <= free_bytes // 2:
        return "stored"
    if full <= int(0.9 * free_bytes) and coded <= free_bytes // 2:
        return "coded"
    return "free"


def random_uniform(n, complex_values=False, seed=0, stream=None):
    """numpy.random.default_rng(seed).random(n) generated on the device
    (edlgt_random_uniform, bit for bit the same PCG64 stream); complex
    output holds the values in the real parts."""
    torch = _torch()
    st = np.random.default_rng(seed).bit_generator.state["state"]
    m64 = (1 << 64) - 1
    state = np.array([st["state"] >> 64, st["state"] & m64, st["inc"] >> 64, st["inc"] & m64], dtype=np.uint64)
    out = torch.empty(int(n), dtype=torch.complex128 if complex_values else torch.float64, device="cuda")
    L.check(L.load().edlgt_random_uniform(L.ptr(out), int(n), L.VALUE_C128 if complex_values else L.VALUE_F64,
                                          state.ctypes.data, L.stream_handle(stream)))
    return out


def expval_terms(basis, tables, placements, psi, stream=None):
    """<psi| strength_p prod_f O_f |psi> for every placement (complex array),
    one device pass over the basis for all of them (edlgt_expval_terms).
    psi: device tensor (complex128 or float64) over the sector basis."""
    tables = np.ascontiguousarray(tables, dtype=np.complex128)
    n_tables, d_max = tables.shape[0], tables.shape[1]
    nops, sites, tabs, stren, _, _ = _pack_placements(placements)
    torch = _torch()
    psi = psi.contiguous()
    vt = L.VALUE_C128 if psi.dtype == torch.complex128 else L.VALUE_F64
    if vt == L.VALUE_F64 and psi.dtype != torch.float64:
        raise ValueError("psi must be complex128 or float64")
    out = np.zeros((len(placements), 2))
    L.check(L.load().edlgt_expval_terms(basis.handle, d_max, n_tables, tables.ctypes.data, len(placements),
                                        nops.ctypes.data, sites.ctypes.data, tabs.ctypes.data,
                                        stren.ctypes.data, L.ptr(psi), vt, out.ctypes.data,
                                        L.stream_handle(stream)))
    return out[:, 0] + 1j * out[:, 1]


class CSRMatrix:
    """Canonical CSR Hamiltonian on the device (rows ascending, columns sorted
    within rows, duplicates summed, explicit zeros kept)."""

    def __init__(self, basis, tables, placements, value_type, stream=None, n_terms=1, matrix_free=False):
        """tables: (n_tables, d, d) complex; placements: list of dicts
        {sites, tables, strength, dagger[, term]}.  With n_terms > 1 every
        placement names its term (0 <= term < n_terms) and the matrix keeps the
        terms apart on a shared pattern (matvec_coef, Server)."""
        lib = L.load()
        tables = np.ascontiguousarray(tables, dtype=np.complex128)
        n_tables, d_max = tables.shape[0], tables.shape[1]
        n_place = len(placements)
        nops, sites, tabs, stren, dag, term = _pack_placements(placements)
        h = C.c_void_p()
        st = L.stream_handle(stream) if stream is not None else None
        self.matrix_free = bool(matrix_free)
        if matrix_free:
            if n_terms != 1:
                raise ValueError("a matrix-free operator has a single term")
            L.check(lib.edlgt_hamiltonian_build_matfree(
                basis.handle, d_max, n_tables, tables.ctypes.data, n_place, nops.ctypes.data,
                sites.ctypes.data, tabs.ctypes.data, stren.ctypes.data, dag.ctypes.data,
                int(value_type), st, C.byref(h)))
        elif n_terms == 1:
            L.check(lib.edlgt_hamiltonian_build(
                basis.handle, d_max, n_tables, tables.ctypes.data, n_place, nops.ctypes.data,
                sites.ctypes.data, tabs.ctypes.data, stren.ctypes.data, dag.ctypes.data,
                int(value_type), st, C.byref(h)))
        else:
            L.check(lib.edlgt_hamiltonian_build_terms(
                basis.handle, d_max, n_tables, tables.ctypes.data, n_place, nops.ctypes.data,
                sites.ctypes.data, tabs.ctypes.data, stren.ctypes.data, dag.ctypes.data,
                term.ctypes.data, int(n_terms), int(value_type), st, C.byref(h)))
        self._h = h
        self.n_terms = int(n_terms)
        n, nnz, vt = C.c_int64(), C.c_int64(), C.c_int32()
        L.check(lib.edlgt_csr_info(h, C.byref(n), C.byref(nnz), C.byref(vt)))
        self.n, self.nnz, self.value_type = int(n.value), int(nnz.value), int(vt.value)
        self.basis = basis

    @classmethod
    def _wrap(cls, h, basis, n_terms=1):
        """A CSRMatrix owning an existing edlgt_csr handle."""
        self = cls.__new__(cls)
        self._h = h
        self.n_terms = int(n_terms)
        self.matrix_free = False
        n, nnz, vt = C.c_int64(), C.c_int64(), C.c_int32()
        L.check(L.load().edlgt_csr_info(h, C.byref(n), C.byref(nnz), C.byref(vt)))
        self.n, self.nnz, self.value_type = int(n.value), int(nnz.value), int(vt.value)
        self.basis = basis
        return self

    @classmethod
    def from_scipy(cls, A, stream=None):
        """A device copy of a square SciPy matrix, duplicates summed
        (edlgt_csr_create): complex128 values when any entry has an imaginary
        part, float64 otherwise."""
        import scipy.sparse as sp
        M = sp.csr_matrix(A)
        M.sum_duplicates()
        if M.shape[0] != M.shape[1]:
            raise ValueError(f"the matrix must be square, got {M.shape}")
        cplx = np.iscomplexobj(M.data) and M.nnz > 0 and np.any(M.data.imag != 0)
        val = np.ascontiguousarray(M.data if cplx else np.real(M.data), dtype=np.complex128 if cplx else np.float64)
        rp = np.ascontiguousarray(M.indptr, dtype=np.int64)
        col = np.ascontiguousarray(M.indices, dtype=np.int32)
        h = C.c_void_p()
        L.check(L.load().edlgt_csr_create(M.shape[0], M.nnz, rp.ctypes.data, col.ctypes.data, val.ctypes.data,
                                          L.VALUE_C128 if cplx else L.VALUE_F64, L.stream_handle(stream),
                                          C.byref(h)))
        return cls._wrap(h, None)

    # SciPy-style sums and scalings (H + c P, as the reference's scripts add a
    # penalty to H.Ham): formed on the host from the stored entries and
    # uploaded as a new device matrix
    def __add__(self, other):
        import scipy.sparse as sp
        if isinstance(other, CSRMatrix):
            other = other.to_scipy()
        if not (sp.issparse(other) or isinstance(other, np.ndarray)):
            return NotImplemented
        return CSRMatrix.from_scipy(self.to_scipy() + sp.csr_matrix(other))

    __radd__ = __add__

    def __mul__(self, a):
        if not np.isscalar(a):
            return NotImplemented
        return CSRMatrix.from_scipy(self.to_scipy() * a)

    __rmul__ = __mul__

    def momentum(self, mb, stream=None):
        """B^H H B on momentum sector mb (edlgt_csr_momentum): a new matrix of
        mb.n_cols rows; this one is left unchanged."""
        h = C.c_void_p()
        L.check(L.load().edlgt_csr_momentum(self._h, mb.handle, L.stream_handle(stream), C.byref(h)))
        out = CSRMatrix._wrap(h, None)
        out.momentum_basis = mb
        return out

    @property
    def handle(self):
        return self._h

    @property
    def is_complex(self):
        return self.value_type == L.VALUE_C128

    @property
    def shape(self):
        return (self.n, self.n)

    def to_host(self):
        """(row_ptr int64, col int32, val) copied to the host."""
        rp = np.empty(self.n + 1, dtype=np.int64)
        col = np.empty(self.nnz, dtype=np.int32)
        val = np.empty(self.nnz, dtype=np.complex128 if self.is_complex else np.float64)
        L.check(L.load().edlgt_csr_copy(self._h, rp.ctypes.data, col.ctypes.data, val.ctypes.data))
        return rp, col, val

    def dense_eigh(self, stream=None):
        """Every eigenpair from the dense matrix on the device
        (edlgt_csr_dense_eigh, rocSOLVER): (evals ascending, evecs tensor
        (N, N), row k = eigenvector k), for up to 32768 states."""
        torch = _torch()
        n = self.n
        if n > 32768:
            raise NotImplementedError("the dense eigensolve takes sectors of at most 32768 states")
        evals = np.zeros(n)
        evecs = torch.empty((n, n), dtype=torch.complex128 if self.is_complex else torch.float64, device="cuda")
        L.check(L.load().edlgt_csr_dense_eigh(self._h, evals.ctypes.data, L.ptr(evecs), L.stream_handle(stream)))
        return evals, evecs

    def compress(self, stream=None):
        """Dictionary-codes the values (edlgt_csr_compress); returns the number
        of distinct (value, term) pairs, 0 when the matrix stays uncoded."""
        nd, cb = C.c_int32(), C.c_int32()
        L.check(L.load().edlgt_csr_compress(self._h, L.stream_handle(stream), C.byref(nd), C.byref(cb)))
        self.n_dict, self.code_bytes = int(nd.value), int(cb.value)
        return self.n_dict

    def release_values(self):
        """Frees the full-width value arrays of a coded matrix (products keep
        working from the codes; to_host can no longer return values)."""
        L.check(L.load().edlgt_csr_release_values(self._h))

    def terms(self):
        """Per-entry term ids (uint8, nnz); all zero for a single-term matrix."""
        t = np.zeros(self.nnz, dtype=np.uint8)
        if self.n_terms > 1:
            L.check(L.load().edlgt_csr_copy_terms(self._h, t.ctypes.data, None))
        return t

    def to_scipy(self, coef=None):
        """The matrix as scipy CSR; for a multi-term matrix, sum_t coef[t] H_t
        (coef defaults to all ones).  Entries of different terms on one
        (row, col) are summed by scipy."""
        import scipy.sparse as sp
        rp, col, val = self.to_host()
        if self.n_terms > 1 and coef is not None:
            val = val * np.asarray(coef, dtype=np.float64)[self.terms()]
        m = sp.csr_matrix((val, col, rp), shape=self.shape)
        if self.n_terms > 1:
            m.sum_duplicates()
        return m

    def toarray(self):
        """Dense host copy (the reference's ham_format="dense" Hamiltonian,
        qmb_hamiltonian.py:218-225), for sectors of at most 16384 states."""
        if self.shape[0] > 16384:
            raise MemoryError(f"a dense copy of a {self.shape[0]}-state Hamiltonian is not made; use to_scipy()")
        return self.to_scipy().toarray()

    def __array__(self, dtype=None, copy=None):
        a = self.toarray()
        return a if dtype is None else a.astype(dtype)

    def vec_type(self, complex_vectors=None):
        cplx = self.is_complex if complex_vectors is None else complex_vectors
        return L.VALUE_C128 if cplx else L.VALUE_F64

    def matvec(self, x, y=None, precision=L.PREC_FP64, stream=None):
        """y = H x for a device tensor x (complex128, or float64 for a real H)."""
        torch = _torch()
        if y is None:
            y = torch.empty_like(x)
        vt = L.VALUE_C128 if x.dtype == torch.complex128 else L.VALUE_F64
        if x.dim() == 1:
            L.check(L.load().edlgt_spmv(self._h, precision, L.ptr(x), L.ptr(y), vt, L.stream_handle(stream)))
        else:
            # x: (n_rhs, N) row-major == N x n_rhs column-major
            L.check(L.load().edlgt_spmm(self._h, precision, x.shape[0], L.ptr(x), L.ptr(y), vt,
                                        L.stream_handle(stream)))
        return y

    def matvec_coef(self, x, coef, y=None, stream=None):
        """Y[j] = (sum_t coef[j, t] H_t) X[j] for device X of shape (n_rhs, N)
        (complex128, or float64 for a real H); coef: (n_rhs, n_terms)."""
        torch = _torch()
        if x.dim() == 1:
            x = x.view(1, -1)
        if y is None:
            y = torch.empty_like(x)
        coef = np.ascontiguousarray(np.broadcast_to(np.asarray(coef, dtype=np.float64),
                                                    (x.shape[0], self.n_terms)))
        vt = L.VALUE_C128 if x.dtype == torch.complex128 else L.VALUE_F64
        L.check(L.load().edlgt_spmm_coef(self._h, x.shape[0], L.ptr(x), L.ptr(y), vt,
                                         coef.ctypes.data, L.stream_handle(stream)))
        return y

    def autotune(self, n_rhs=1, precision=L.PREC_FP64, complex_vectors=None, stream=None):
        """Times the SpMM launch shapes on scratch vectors and keeps the fastest.
        Returns (lanes_per_row, workgroups, ms)."""
        torch = _torch()
        vt = self.vec_type(complex_vectors)
        dt = torch.complex128 if vt == L.VALUE_C128 else torch.float64
        X = torch.ones((n_rhs, self.n), dtype=dt, device="cuda")
        Y = torch.empty_like(X)
        g, grid, ms = C.c_int32(), C.c_int64(), C.c_float()
        L.check(L.load().edlgt_csr_autotune(self._h, precision, vt, n_rhs, L.ptr(X), L.ptr(Y),
                                            L.stream_handle(stream), C.byref(g), C.byref(grid),
                                            C.byref(ms)))
        return int(g.value), int(grid.value), float(ms.value)

    def set_launch(self, lanes_per_row=0, grid=0):
        L.check(L.load().edlgt_csr_set_launch(self._h, lanes_per_row, grid))

    def lanczos(self, n_eigs, tol=1e-12, max_basis=0, max_iter=100000, precision=L.PREC_FP64,
                v0=None, complex_vectors=None, want_vectors=True, stream=None):
        """Lowest eigenpairs: (evals ascending, evecs tensor (k, N) or None, resid, matvecs).
        max_basis 0: eigsh's default basis size max(2k + 1, 20), at least 16
        blocks of min(k, 4) vectors when k > 1."""
        torch = _torch()
        vt = self.vec_type(complex_vectors)
        if v0 is not None:
            if hasattr(v0, "data_ptr"):  # device tensor: used in place
                want = torch.complex128 if vt == L.VALUE_C128 else torch.float64
                if v0.dtype != want or not v0.is_cuda:
                    raise ValueError(f"device v0 must be a {want} CUDA tensor")
                v0 = v0.contiguous()
            else:
                v0 = np.ascontiguousarray(v0, dtype=np.complex128 if vt == L.VALUE_C128 else np.float64)
            if (v0.numel() if hasattr(v0, "numel") else v0.size) != self.n:
                raise ValueError("v0 has the wrong length")
        evals = np.zeros(n_eigs)
        resid = np.zeros(n_eigs)
        mvs = np.zeros(2, dtype=np.int64)
        evecs = None
        if want_vectors:
            evecs = torch.empty((n_eigs, self.n), dtype=torch.complex128 if vt == L.VALUE_C128 else torch.float64,
                                device="cuda")
        L.check(L.load().edlgt_lanczos(self._h, n_eigs, tol, max_basis, max_iter, precision, vt,
                                       L.ptr(v0), evals.ctypes.data, L.ptr(evecs), resid.ctypes.data,
                                       mvs.ctypes.data, L.stream_handle(stream)))
        return evals, evecs, resid, mvs

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.load().edlgt_csr_destroy(h)
            except Exception:
                pass
            self._h = None


class Server:
    """Served lowest-eigenvalue solves with continuous batching on a
    multi-term matrix (edlgt_server_*).  Each request is a coefficient vector
    c (H(c) = sum_t c[t] H_t) and a seed; up to max_slots requests advance
    together, one fused SpMM per step, and a finished request frees its slot
    at once for the next one.  Slots are allocated on first use
    (slot_bytes each).  With warm_cache = 0 a request's result does not depend
    on which other requests share the batch (bitwise); with warm_cache = C the
    server starts new requests from the stored Ritz vectors of the nearest of
    up to C finished couplings (results equal to tol)."""

    def __init__(self, csr, n_eigs=1, tol=1e-10, max_basis=32, max_matvecs=100000,
                 max_slots=8, complex_vectors=None, warm_cache=0, stream=None):
        lib = L.load()
        self.csr = csr
        self.n_eigs = int(n_eigs)
        self.max_slots = int(max_slots)
        self._stream = L.stream_handle(stream)
        vt = csr.vec_type(complex_vectors)
        s = C.c_void_p()
        L.check(lib.edlgt_server_create(csr.handle, vt, self.max_slots, self.n_eigs, float(tol),
                                        int(max_basis), int(max_matvecs), int(warm_cache),
                                        self._stream, C.byref(s)))
        self._s = s
        self._ids = np.zeros(self.max_slots, dtype=np.int64)
        self._ev = np.zeros((self.max_slots, self.n_eigs))
        self._es = np.zeros((self.max_slots, self.n_eigs))
        self._mv = np.zeros(self.max_slots, dtype=np.int64)
        self._cv = np.zeros(self.max_slots, dtype=np.int32)

    def submit(self, request_id, coef, seed):
        """Puts the request into a free slot; returns False when all are busy."""
        c = np.zeros(self.csr.n_terms)
        c[:] = np.asarray(coef, dtype=np.float64)
        slot = C.c_int32()
        L.check(L.load().edlgt_server_submit(self._s, int(request_id), c.ctypes.data,
                                             C.c_uint64(int(seed)), C.byref(slot)))
        return slot.value >= 0

    def info(self):
        """dict(active, allocated, max_slots, slot_bytes, shared_bytes)."""
        a, al, mx = C.c_int32(), C.c_int32(), C.c_int32()
        sb, wb = C.c_int64(), C.c_int64()
        L.check(L.load().edlgt_server_info(self._s, C.byref(a), C.byref(al), C.byref(mx), C.byref(sb),
                                           C.byref(wb)))
        return dict(active=a.value, allocated=al.value, max_slots=mx.value, slot_bytes=sb.value,
                    shared_bytes=wb.value)

    @property
    def active(self):
        return self.info()["active"]

    def trim(self):
        """Frees the memory of every idle slot."""
        L.check(L.load().edlgt_server_trim(self._s))

    def step(self):
        """One fused expansion step; returns the requests that finished as a
        list of (request_id, evals, resid_est, matvecs, converged)."""
        n = C.c_int32()
        L.check(L.load().edlgt_server_step(self._s, self.max_slots, self._ids.ctypes.data,
                                           self._ev.ctypes.data, self._es.ctypes.data,
                                           self._mv.ctypes.data, self._cv.ctypes.data, C.byref(n)))
        return [(int(self._ids[i]), self._ev[i].copy(), self._es[i].copy(), int(self._mv[i]),
                 bool(self._cv[i])) for i in range(n.value)]

    def solve_all(self, requests):
        """Runs a list of (request_id, coef, seed) to completion, admitting in
        order as slots free up; returns {request_id: (evals, resid, matvecs, converged)}."""
        out, queue = {}, list(requests)
        while queue or self.active:
            while queue and self.submit(*queue[0]):
                queue.pop(0)
            for rid, ev, es, mv, cv in self.step():
                out[rid] = (ev, es, mv, cv)
        return out

    def __del__(self):
        s = getattr(self, "_s", None)
        if s is not None and s.value:
            try:
                L.load().edlgt_server_destroy(s)
            except Exception:
                pass
            self._s = None


class KrylovSession:
    """Krylov exp(-iH dt) stepping state kept on the device between calls
    (edlgt_krylov_create / run / destroy): the m+1 basis vectors and the
    marginal buffer are allocated once, so a trajectory evolved in chunks does
    not re-allocate them per chunk."""

    def __init__(self, csr, basis=None, precision=L.PREC_FP64, m=30, reorth=True, d_max=0, stream=None):
        k = C.c_void_p()
        L.check(L.load().edlgt_krylov_create(csr.handle, None if basis is None else basis.handle,
                                             int(precision), int(m), int(reorth), int(d_max or 0),
                                             L.stream_handle(stream), C.byref(k)))
        self._k = k
        self.csr, self.basis = csr, basis
        self.key = (id(csr), None if basis is None else id(basis), int(precision), int(m), bool(reorth),
                    int(d_max or 0))

    def run(self, psi, dt, n_steps, marginals=None, err=None, history=None):
        """psi <- exp(-i H dt)^n_steps psi (device complex128, in place); per-step
        marginals (host, (n_steps+1, n_sites, d_max)), error estimates (host) and
        states (device, (n_steps+1, N)) when given."""
        L.check(L.load().edlgt_krylov_run(self._k, L.ptr(psi), float(dt), int(n_steps), L.ptr(marginals),
                                          L.ptr(err), L.ptr(history)))

    def close(self):
        k = getattr(self, "_k", None)
        if k is not None and k.value:
            try:
                L.load().edlgt_krylov_destroy(k)
            except Exception:
                pass
        self._k = None

    def __del__(self):
        self.close()


class MomentumBasis:
    """Projector B onto momentum k of the block translations by unit_cell
    along each axis (edlgt_momentum_*; get_momentum_basis,
    translational_sym.py:1307-1359).  n_cols columns, one per compatible
    translation orbit; to_scipy() gives B, expand() / project() apply B and B^H
    to device vectors."""

    def __init__(self, basis, lvals, unit_cell, k_vals, stream=None, tc_map=None, tc_sign=None):
        """tc_map / tc_sign given: the translation-combined symmetry of a 1D
        ring (shift by one site after the local involution tc_map, odd powers
        signed by prod tc_sign over the configuration), unit_cell ignored."""
        lv = np.ascontiguousarray(lvals, dtype=np.int32)
        uc = np.ascontiguousarray(unit_cell, dtype=np.int32)
        kv = np.ascontiguousarray(np.atleast_1d(k_vals), dtype=np.int32)
        m = C.c_void_p()
        if tc_map is not None:
            cm = np.ascontiguousarray(tc_map, dtype=np.int32)
            cs = np.ascontiguousarray(tc_sign, dtype=np.float64)
            if lv.size != 1 or kv.size != 1 or cs.size != cm.size:
                raise ValueError("TC momentum sectors are 1D: one k and a sign per local state")
            L.check(L.load().edlgt_momentum_build_tc(basis.handle, int(cm.size), cm.ctypes.data, cs.ctypes.data,
                                                     int(kv[0]), L.stream_handle(stream), C.byref(m)))
        else:
            if not (lv.size == uc.size == kv.size):
                raise ValueError("lvals, unit_cell and k_vals need one entry per axis")
            L.check(L.load().edlgt_momentum_build(basis.handle, int(lv.size), lv.ctypes.data, uc.ctypes.data,
                                                  kv.ctypes.data, L.stream_handle(stream), C.byref(m)))
        self._m = m
        self.basis = basis
        self.k_vals, self.unit_cell = kv.copy(), uc.copy()
        nr, nc, nnz, cx = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        L.check(L.load().edlgt_momentum_info(m, C.byref(nr), C.byref(nc), C.byref(nnz), C.byref(cx)))
        self.n_rows, self.n_cols, self.nnz, self.is_complex = int(nr.value), int(nc.value), int(nnz.value), bool(cx.value)

    @property
    def handle(self):
        return self._m

    def to_scipy(self):
        import scipy.sparse as sp
        cp = np.zeros(self.n_cols + 1, dtype=np.int64)
        ri = np.zeros(max(self.nnz, 1), dtype=np.int32)
        dv = np.zeros((max(self.nnz, 1), 2))
        L.check(L.load().edlgt_momentum_copy(self._m, cp.ctypes.data, ri.ctypes.data, dv.ctypes.data))
        v = dv[: self.nnz, 0] + 1j * dv[: self.nnz, 1]
        return sp.csc_matrix((v, ri[: self.nnz], cp), shape=(self.n_rows, self.n_cols))

    def _apply(self, forward, x, stream=None):
        torch = _torch()
        x = x.to(device="cuda", dtype=torch.complex128).contiguous()
        if x.numel() != (self.n_cols if forward else self.n_rows):
            raise ValueError("vector length does not match the momentum basis")
        y = torch.empty(self.n_rows if forward else self.n_cols, dtype=torch.complex128, device="cuda")
        L.check(L.load().edlgt_momentum_apply(self._m, 1 if forward else 0, L.ptr(x), L.ptr(y),
                                              L.stream_handle(stream)))
        return y

    def expand(self, psi_k, stream=None):
        """B psi_k: a momentum-sector vector in the sector basis."""
        return self._apply(True, psi_k, stream)

    def project(self, psi, stream=None):
        """B^H psi."""
        return self._apply(False, psi, stream)

    def __del__(self):
        m = getattr(self, "_m", None)
        if m is not None and m.value:
            try:
                L.load().edlgt_momentum_destroy(m)
            except Exception:
                pass
            self._m = None
