"""Python handles around the device objects of libedlgt_hip.

SectorBasis  -- sorted symmetry-sector basis (keys + configurations) on the GPU
CSRMatrix    -- canonical CSR Hamiltonian on the GPU, with SpMV/SpMM and Lanczos
"""
import ctypes as C

import numpy as np

from . import _lib as L


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
        return self._configs

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


class CSRMatrix:
    """Canonical CSR Hamiltonian on the device (rows ascending, columns sorted
    within rows, duplicates summed, explicit zeros kept)."""

    def __init__(self, basis, tables, placements, value_type, stream=None):
        """tables: (n_tables, d, d) complex; placements: list of dicts
        {sites, tables, strength, dagger}."""
        lib = L.load()
        tables = np.ascontiguousarray(tables, dtype=np.complex128)
        n_tables, d_max = tables.shape[0], tables.shape[1]
        n_place = len(placements)
        nops = np.zeros(n_place, dtype=np.int32)
        sites = np.zeros((n_place, 8), dtype=np.int32)
        tabs = np.zeros((n_place, 8), dtype=np.int32)
        stren = np.zeros((n_place, 2), dtype=np.float64)
        dag = np.zeros(n_place, dtype=np.int32)
        for p, pl in enumerate(placements):
            k = len(pl["sites"])
            nops[p] = k
            sites[p, :k] = pl["sites"]
            tabs[p, :k] = pl["tables"]
            s = complex(pl["strength"])
            stren[p] = (s.real, s.imag)
            dag[p] = 1 if pl.get("dagger") else 0
        h = C.c_void_p()
        st = L.stream_handle(stream) if stream is not None else None
        L.check(lib.edlgt_hamiltonian_build(
            basis.handle, d_max, n_tables, tables.ctypes.data, n_place, nops.ctypes.data,
            sites.ctypes.data, tabs.ctypes.data, stren.ctypes.data, dag.ctypes.data,
            int(value_type), st, C.byref(h)))
        self._h = h
        n, nnz, vt = C.c_int64(), C.c_int64(), C.c_int32()
        L.check(lib.edlgt_csr_info(h, C.byref(n), C.byref(nnz), C.byref(vt)))
        self.n, self.nnz, self.value_type = int(n.value), int(nnz.value), int(vt.value)
        self.basis = basis

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

    def to_scipy(self):
        import scipy.sparse as sp
        rp, col, val = self.to_host()
        return sp.csr_matrix((val, col, rp), shape=self.shape)

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

    def lanczos(self, n_eigs, tol=1e-12, max_basis=64, max_iter=100000, precision=L.PREC_FP64,
                v0=None, complex_vectors=None, want_vectors=True, stream=None):
        """Lowest eigenpairs: (evals ascending, evecs tensor (k, N) or None, resid, matvecs)."""
        torch = _torch()
        vt = self.vec_type(complex_vectors)
        if v0 is not None:
            v0 = np.ascontiguousarray(v0, dtype=np.complex128 if vt == L.VALUE_C128 else np.float64)
            if v0.size != self.n:
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
