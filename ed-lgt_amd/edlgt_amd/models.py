"""Model API of the engine, mirroring edlgt's QuantumModel / QED_Model /
QMB_hamiltonian / QMB_state (edlgt/models/quantum_model.py,
edlgt/models/QED_model.py, edlgt/modeling/qmb_hamiltonian.py,
edlgt/modeling/qmb_state.py): same constructor arguments, method names,
attributes and per-site energy convention.  Sector enumeration, Hamiltonian
assembly, eigensolves and time evolution run on the GPU through
libedlgt_hip; only the tiny per-site operator tables are built on the host.
"""
import logging
from fractions import Fraction
from math import prod

import numpy as np

from . import _lib as L
from .device import CSRMatrix, SectorBasis
from .lattice import Lattice
from . import operators as opx

logger = logging.getLogger(__name__)

__all__ = ["QuantumModel", "QED_Model", "SU2_Model", "Zn_Model", "QMB_hamiltonian", "QMB_state"]


class QMB_state:
    """State vector wrapper (edlgt/modeling/qmb_state.py:41-77).  `psi` is the
    host copy (numpy), `psi_device` the device tensor; constructed from either
    (a host array is copied to the device once)."""

    def __init__(self, psi, lvals=None, loc_dims=None, symmetry_sector=True, debug_mode=False):
        if hasattr(psi, "data_ptr"):
            self.psi_device = psi
            self._host = None
        else:
            import torch
            host = np.asarray(psi)
            if host.ndim != 1 or host.size == 0:
                raise ValueError(f"psi must be a non-empty 1D array, got shape {host.shape}")
            if not np.iscomplexobj(host):
                host = host.astype(np.float64)
            self.psi_device = torch.as_tensor(np.ascontiguousarray(host), device="cuda")
            self._host = None
        self.lvals = lvals
        self.loc_dims = loc_dims
        self.symmetry_sector = symmetry_sector
        self.debug_mode = debug_mode

    @property
    def psi(self):
        if self._host is None:
            self._host = self.psi_device.detach().cpu().numpy()
        return self._host

    def get_norm(self):
        import torch
        return float(torch.linalg.vector_norm(self.psi_device).item())

    def normalize(self, threshold=1e-14):
        """Rescales the state to unit norm when its norm differs from 1 by more
        than threshold; returns the norm before (qmb_state.py:79-95)."""
        norm = self.get_norm()
        if abs(norm - 1.0) > threshold:
            self.psi_device = self.psi_device / norm
            self._host = None
        return norm

    def truncate(self, threshold=1e-14):
        """Host copy of the amplitudes with |psi_i| <= threshold set to zero
        (qmb_state.py:97-110, truncation :466-482); the state is unchanged."""
        if not np.isscalar(threshold) or threshold < 0:
            raise ValueError(f"threshold must be a non-negative scalar, not {threshold!r}")
        psi = self.psi
        return np.where(np.abs(psi) > threshold, psi, 0)

    def expectation_value(self, operator, component="real"):
        """<psi|O|psi> (qmb_state.py:112-157) for O given as (rows, cols, values)
        triplets (summed like exp_val_data, duplicates included), a dense
        matrix, a SciPy sparse matrix, or a CSRMatrix on the device (one device
        product).  The first three are applied to the host copy of psi, as the
        reference does; component: "real" or "imag"."""
        import scipy.sparse as sps
        if component not in ("real", "imag"):
            raise ValueError(f"component must be 'real' or 'imag', not {component!r}")
        if isinstance(operator, CSRMatrix):
            import torch
            psi = self.psi_device
            if operator.n != psi.numel():
                raise ValueError(f"quantum state dim {psi.numel()} is NOT the operator dim {operator.n}")
            if not operator.is_complex and psi.is_complex():
                y = torch.complex(operator.matvec(psi.real.contiguous()), operator.matvec(psi.imag.contiguous()))
            else:
                y = operator.matvec(psi.contiguous())
            val = complex(torch.vdot(psi, y).item())
        elif isinstance(operator, tuple) and len(operator) == 3:
            r, c, v = (np.asarray(a) for a in operator)
            psi = self.psi
            val = complex(np.sum(np.conj(psi[r]) * v * psi[c]))
        elif isinstance(operator, np.ndarray) or sps.issparse(operator):
            psi = self.psi
            if operator.ndim != 2 or operator.shape[0] != operator.shape[1] or operator.shape[0] != psi.shape[0]:
                raise ValueError(f"quantum state dim {psi.shape[0]} is NOT the operator dim {operator.shape[0]}")
            val = complex(np.vdot(psi, sps.csr_matrix(operator) @ psi))
        else:
            raise TypeError("Operator must be provided as a tuple (row_list, col_list, value_list), "
                            "a dense matrix (np.ndarray), or a sparse matrix (scipy sparse matrix).")
        return val.real if component == "real" else val.imag

    def participation_renyi_entropy(self, alpha=2):
        """PE_alpha = log(sum_i |psi_i|^(2 alpha) + 1e-16) / (1 - alpha) in the
        sector basis (qmb_state.py:357-391), reduced on the device."""
        import torch
        if alpha <= 0 or alpha == 1:
            raise ValueError(f"alpha must be > 0 and != 1. Got alpha={alpha}.")
        p = torch.abs(self.psi_device).double() ** 2
        ipr = float(torch.sum(p ** float(alpha)).item())
        return float(np.log(ipr + 1e-16) / (1.0 - float(alpha)))

    def stabilizer_renyi_entropy(self, sector_configs, prob_threshold=1e-2):
        """Stabilizer Renyi-2 entropy on the support holding >= 1 -
        prob_threshold of the weight (qmb_state.py:393-465), on the device.
        sector_configs: the model's sector_configs (or its SectorBasis), whose
        device basis the state is expanded in."""
        from .device import SectorBasis, basis_of_configs
        basis = sector_configs if isinstance(sector_configs, SectorBasis) else basis_of_configs(sector_configs)
        if basis is None:
            raise ValueError("sector_configs must be the model's sector_configs (the device basis of the state)")
        sre, K, disc, n_str = basis.sre2(self.psi_device, prob_threshold)
        logger.info(f"Support size = {K}, discarded weight = {disc:.3e}, {n_str} X-strings; SRE2 = {sre:.12f}")
        return sre

    def entanglement_entropy(self, keep_indices, partitions_dict):
        """Von Neumann entropy (base 2, Schmidt weights > 1e-10) of the sites
        keep_indices (qmb_state.py:245-291), from the subsystem/environment
        maps of QuantumModel._get_partition.  The coefficient matrix
        psi[subsystem config, environment config] of a symmetry sector is block
        diagonal up to permutations; its squared singular values are computed
        on the device block by block (schmidt_spectrum) -- the spectrum of the
        reference's one dense SVD."""
        part = partitions_dict[tuple(sorted(keep_indices))]
        if min(part["subsys_dim"], part["env_dim"]) == 1:
            return 0.0  # rank one, as in the reference
        lam = schmidt_spectrum(self.psi_device, part)
        lam = lam[lam > 1e-10]
        return float(-np.sum(lam * np.log2(lam)))

    def reduced_density_matrix(self, keep_indices, partitions_dict):
        """Dense reduced density matrix of the sites keep_indices, psi_matrix
        psi_matrix^H over the sector's subsystem configurations
        (qmb_state.py:211-243): its diagonal blocks are formed on the device
        and placed on the host."""
        part = partitions_dict[tuple(sorted(keep_indices))]
        lay = schmidt_layout(part, "subsys")
        G = schmidt_gram(self.psi_device, lay).cpu().numpy()
        rho = np.zeros((part["subsys_dim"], part["subsys_dim"]), dtype=np.complex128)
        for b, rows in enumerate(lay["row_ids"]):
            d, _, _, goff = lay["blocks"][b]
            rho[np.ix_(rows, rows)] = G[goff:goff + d * d].reshape(d, d)
        return rho

    def get_state_configurations(self, threshold, sector_configs):
        """(configurations, probabilities) of the basis states with |psi|^2 >
        threshold, largest first (qmb_state.py get_state_configurations)."""
        p = np.abs(self.psi) ** 2
        idx = np.nonzero(p > threshold)[0]
        idx = idx[np.argsort(-p[idx], kind="stable")]
        for i in idx:
            logger.info(f"{sector_configs[i]}  {p[i]:.6f}")
        return sector_configs[idx], p[idx]


def schmidt_layout(part, rows="small"):
    """Block layout of psi[subsystem, environment] for the device Schmidt
    kernels (edlgt_schmidt_gram), cached in the partition dict.  The blocks
    are the connected components of the bipartite graph of (subsystem,
    environment) pairs present in the sector; block b is a dense d x c
    row-major matrix whose rows are its subsystem configurations
    (rows="subsys", for the reduced density matrix) or its smaller side
    (rows="small", for the spectrum), ordered by configuration index.
    Returns dict(blocks=int64 [nb, 4] (d, c, moff, goff), pos_host=int64
    element position of every basis state in the packed blocks (copied to the
    device on first use), row_ids=subsystem indices of each block's rows
    (rows="subsys"), n_vals, m_elems, g_elems)."""
    key = ("schmidt", rows)
    if key in part:
        return part[key]
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    a, e = np.asarray(part["subsys_map"], np.int64), np.asarray(part["env_map"], np.int64)
    na, ne = int(part["subsys_dim"]), int(part["env_dim"])
    g = sp.coo_matrix((np.ones(a.size), (a, na + e)), shape=(na + ne, na + ne))
    nc, lab = connected_components(g, directed=False)
    comp = lab[a].astype(np.int64)  # component of every basis state
    # rank of each state's subsystem (environment) configuration inside its component
    ka, ia = np.unique(comp * na + a, return_inverse=True)
    ke, ie = np.unique(comp * ne + e, return_inverse=True)
    ia, ie = ia.ravel().astype(np.int64), ie.ravel().astype(np.int64)
    ca, ce = ka // na, ke // ne  # component of each unique key (sorted)
    fa = np.searchsorted(ca, np.arange(nc + 1))
    fe = np.searchsorted(ce, np.arange(nc + 1))
    ra, re_ = ia - fa[comp], ie - fe[comp]
    da, de = np.diff(fa), np.diff(fe)
    used = np.flatnonzero(da > 0)  # components with basis states
    if rows == "subsys":
        sub_rows = np.ones(nc, bool)
    elif rows == "small":
        sub_rows = da <= de
    else:
        raise ValueError(f"rows must be 'small' or 'subsys', not {rows!r}")
    d = np.where(sub_rows, da, de)[used]
    c = np.where(sub_rows, de, da)[used]
    moff = np.concatenate([[0], np.cumsum(d * c)[:-1]])
    goff = np.concatenate([[0], np.cumsum(d * d)[:-1]])
    bidx = np.full(nc, -1, np.int64)
    bidx[used] = np.arange(used.size)
    b = bidx[comp]
    r = np.where(sub_rows[comp], ra, re_)
    col = np.where(sub_rows[comp], re_, ra)
    pos = moff[b] + r * c[b] + col
    lay = dict(blocks=np.ascontiguousarray(np.stack([d, c, moff, goff], axis=1).astype(np.int64)),
               pos_host=np.ascontiguousarray(pos, dtype=np.int64),
               n_vals=int(d.sum()), m_elems=int((d * c).sum()), g_elems=int((d * d).sum()),
               row_ids=[ka[fa[k]:fa[k + 1]] - k * na for k in used] if rows == "subsys" else None)
    part[key] = lay
    return lay


def schmidt_gram(psi_device, lay):
    """Device complex128 array of the Gram blocks G_b = M_b M_b^H of the
    layout (edlgt_schmidt_gram)."""
    import torch
    psi = psi_device.reshape(-1)
    if psi.dtype not in (torch.complex128, torch.float64):
        psi = psi.to(torch.complex128)
    psi = psi.contiguous()
    if psi.numel() != lay["pos_host"].size:
        raise ValueError(f"state has {psi.numel()} entries, the partition {lay['pos_host'].size}")
    if "pos" not in lay:
        lay["pos"] = torch.as_tensor(lay["pos_host"], device=psi.device)
    vt = L.VALUE_C128 if psi.dtype == torch.complex128 else L.VALUE_F64
    G = torch.empty(max(lay["g_elems"], 1), dtype=torch.complex128, device=psi.device)
    blocks = lay["blocks"]
    L.check(L.load().edlgt_schmidt_gram(L.ptr(psi), vt, psi.numel(), L.ptr(lay["pos"]), blocks.shape[0],
                                        L.ptr(blocks), L.ptr(G), G.numel(), L.stream_handle()))
    return G


def schmidt_spectrum(psi_device, part):
    """Squared Schmidt values of the state over the partition (every block's
    Gram eigenvalues, edlgt_hermitian_eigvals; unsorted)."""
    lay = schmidt_layout(part, "small")
    G = schmidt_gram(psi_device, lay)
    ev = np.empty(max(lay["n_vals"], 1))
    blocks = lay["blocks"]
    L.check(L.load().edlgt_hermitian_eigvals(L.ptr(G), G.numel(), blocks.shape[0], L.ptr(blocks), L.ptr(ev),
                                             ev.size, L.stream_handle()))
    return ev[:lay["n_vals"]]


def get_entropy_partition(lvals, option="half"):
    """Sites of the standard half bipartition (qmb_hamiltonian.py:635-664): the
    first floor(Lx/2) sites of a chain, or of each of the first two rows."""
    if option != "half":
        raise NotImplementedError("for the moment it works only for option=half")
    half = list(range(int(lvals[0] / 2)))
    return half if len(lvals) == 1 else half + [lvals[0] + i for i in half]


def diagonalize_density_matrix(rho):
    """(eigenvalues ascending, eigenvectors) of a Hermitian density matrix
    (qmb_state.py:526-546)."""
    if hasattr(rho, "toarray"):
        rho = rho.toarray()
    return np.linalg.eigh(np.asarray(rho))


def get_projector_for_efficient_density_matrix(rho_eigvals, rho_eigvecs, threshold):
    """Columns of the eigenvectors whose eigenvalues exceed threshold, largest
    first; the threshold divided by 10 until at least two pass
    (qmb_state.py:549-592)."""
    order = np.argsort(rho_eigvals)[::-1]
    vals = np.asarray(rho_eigvals)[order]
    vecs = np.asarray(rho_eigvecs)[:, order]
    n_cols = int(np.sum(vals > threshold))
    while n_cols < 2:
        threshold /= 10
        n_cols = int(np.sum(vals > threshold))
    logger.info(f"SIGNIFICANT EIGENVALUES {n_cols} with threshold {threshold}")
    return vecs[:, :n_cols].astype(complex)


def truncation(array, threshold=1e-14):
    """Entries with |a| <= threshold set to zero (qmb_state.py:466-482)."""
    a = np.asarray(array)
    return np.where(np.abs(a) > threshold, a, 0)


def get_norm(psi):
    """Euclidean norm of a complex vector (qmb_state.py:484-502)."""
    return float(np.sqrt(np.sum(np.abs(np.asarray(psi)) ** 2)))


def get_sorted_indices(data):
    """Indices ordering data by descending magnitude, ties by descending real
    part (qmb_state.py:504-523: lexsort on (real, abs), reversed)."""
    data = np.asarray(data)
    return np.lexsort((np.real(data), np.abs(data)))[::-1]


def exact_time_evolution(tline, eigenvalues, eigenvectors, initial_state):
    """psi(t) = sum_i e^{-i E_i t} <E_i|psi0> E_i for every t of tline, rows
    of the result; eigenvectors by columns (qmb_hamiltonian.py:792-836)."""
    V = np.asarray(eigenvectors)
    c = V.conj().T @ np.asarray(initial_state, dtype=np.complex128)
    ph = np.exp(-1j * np.outer(np.asarray(tline, dtype=float), np.asarray(eigenvalues, dtype=float)))
    return (ph * c[None, :]) @ V.T


def compute_overlap_with_eigenstates(eigenvectors, state):
    """<E_i|state> for the eigenvectors given by columns
    (qmb_hamiltonian.py:838-861)."""
    return np.asarray(eigenvectors).conj().T @ np.asarray(state, dtype=np.complex128)


def build_energy_block_ids(evals, tol=1e-10):
    """(block id per eigenvalue, number of blocks): consecutive sorted
    eigenvalues closer than tol share a degenerate block
    (quantum_model.py:1451-1478)."""
    E = np.real_if_close(np.asarray(evals))
    if E.size == 0:
        return np.zeros(0, dtype=np.int32), 0
    block = np.concatenate([[0], np.cumsum(np.abs(np.diff(E)) >= tol)]).astype(np.int32)
    return block, int(block[-1]) + 1


def apply_projection(projector, operator):
    """P^dagger O P in the dense or sparse family of the inputs
    (quantum_model.py:1411-1448)."""
    import scipy.sparse as sps
    if projector.shape[0] != operator.shape[0]:
        raise ValueError(f"Projector and operator have incompatible shapes {projector.shape} {operator.shape}")
    if sps.issparse(projector) and sps.issparse(operator):
        return (projector.conj().T @ operator @ projector).tocsr()
    if not sps.issparse(projector) and not sps.issparse(operator):
        P = np.asarray(projector)
        return P.conj().T @ np.asarray(operator) @ P
    raise TypeError("projector and operator must both be dense or both be sparse")


def format_loc_dims(loc_dims, lvals, pad=3):
    """Logs the local dimensions on the lattice's geometry: one line per row
    of x; in 3D the z layers of a row side by side (quantum_model.py:
    1481-1523)."""
    arr = np.asarray(loc_dims)

    def row(vals):
        return "[" + " ".join(f"{v:>{pad}d}" for v in vals) + "]"

    if len(lvals) == 1:
        logger.info(row(arr.reshape(lvals[0])))
    elif len(lvals) == 2:
        for r in arr.reshape(lvals[1], lvals[0]):
            logger.info(row(r))
    elif len(lvals) == 3:
        g = arr.reshape(lvals[2], lvals[1], lvals[0])  # [z, y, x]
        for yy in range(lvals[1]):
            logger.info("   ".join(row(g[zz, yy]) for zz in range(lvals[2])))


def mixed_exp_val_data(psi1, psi2, row_list, col_list, value_list):
    """<psi1|O|psi2> of an operator given as triplets, duplicates summed
    (qmb_state.py:657-679)."""
    psi1, psi2 = np.asarray(psi1), np.asarray(psi2)
    r, c = np.asarray(row_list), np.asarray(col_list)
    return complex(np.sum(np.conj(psi1[r]) * np.asarray(value_list) * psi2[c]))


class QMB_hamiltonian:
    """Hamiltonian container (edlgt/modeling/qmb_hamiltonian.py:24-426) whose
    matrix lives on the GPU as canonical CSR."""

    def __init__(self, lvals, size):
        self.lvals = list(lvals)
        self.shape = (size, size)
        self.Ham = None
        self.Ham_type = None

    def print_energy(self, en_state):
        """qmb_hamiltonian.py:603-612."""
        logger.info("====================================================")
        logger.info(f"{en_state} ENERGY: {round(self.Nenergies[en_state], 16)}")

    @property
    def Ham(self):
        return self.__dict__.get("_Ham")

    @Ham.setter
    def Ham(self, value):
        """Any new matrix (also H.Ham = H.Ham + c P, as the reference's scripts
        write it) drops the levels solved for the previous one."""
        self._Ham = value
        self._all_pairs = None
        if value is not None and hasattr(value, "nnz"):
            self.nnz = value.nnz

    def set_matrix(self, csr, fmt):
        self.Ham = csr
        self.Ham_type = fmt

    def add_term(self, term):
        """Accumulates a contribution given as (row_list, col_list,
        value_list) triplets, a dense matrix (its nonzero entries) or a SciPy
        sparse matrix (qmb_hamiltonian.py:155-171); build() assembles them."""
        import scipy.sparse as sps
        if isinstance(term, tuple) and len(term) == 3:
            r, c, v = (np.asarray(a) for a in term)
        elif isinstance(term, np.ndarray):
            r, c = np.nonzero(term)
            v = term[r, c]
        elif sps.issparse(term):
            C_ = sps.coo_matrix(term)
            r, c, v = C_.row, C_.col, C_.data
        else:
            raise TypeError("Unsupported term type: expect (dense/sparse) or (row_list, col_list, value_list) ")
        r, c, v = np.ravel(r).astype(np.int64), np.ravel(c).astype(np.int64), np.ravel(v)
        if not (r.size == c.size == v.size):
            raise ValueError("row_list, col_list and value_list must have the same length")
        if not hasattr(self, "_pending") or self._pending is None:
            self._pending = []
        self._pending.append((r, c, v))

    def build(self, format):
        """The accumulated triplets as SciPy's coalesced CSR (duplicates summed,
        explicit zeros kept), uploaded as the device matrix
        (qmb_hamiltonian.py:173-225); every format is served by the device
        matrix, the label selects the time-evolution path."""
        import scipy.sparse as sps
        from .device import CSRMatrix
        if format not in ("dense", "sparse", "linear"):
            raise ValueError(f"Invalid format: {format}. Use 'dense', 'sparse', or 'linear'.")
        parts = getattr(self, "_pending", None) or []
        n = self.shape[0]
        if parts:
            r = np.concatenate([p[0] for p in parts])
            c = np.concatenate([p[1] for p in parts])
            v = np.concatenate([np.asarray(p[2], dtype=np.complex128) for p in parts])
        else:
            r = c = np.zeros(0, dtype=np.int64)
            v = np.zeros(0, dtype=np.complex128)
        M = sps.csr_matrix((v, (r, c)), shape=(n, n))
        logger.info(f"Construct {format} Hamiltonian")
        self.set_matrix(CSRMatrix.from_scipy(M), format)
        self.get_sparsity()

    def _triplets(self):
        parts = getattr(self, "_pending", None)
        if self.Ham is None and parts:
            return (np.concatenate([p[0] for p in parts]).astype(np.int32),
                    np.concatenate([p[1] for p in parts]).astype(np.int32),
                    np.concatenate([np.asarray(p[2]) for p in parts]))
        H = self.Ham
        if hasattr(H, "to_scipy"):
            H = H.to_scipy()
        C = H.tocoo()
        return C.row.astype(np.int32), C.col.astype(np.int32), C.data

    @property
    def row_list(self):
        """Rows of the stored entries (the reference keeps its triplets here)."""
        return self._triplets()[0]

    @property
    def col_list(self):
        return self._triplets()[1]

    @property
    def value_list(self):
        return self._triplets()[2]

    def convert_hamiltonian(self, format):
        """qmb_hamiltonian.py:227-258.  The device matrix serves every format
        (all products run on the device); the label selects the time-evolution
        path as in the reference ("dense": exact, from every eigenpair)."""
        if format not in ("dense", "sparse", "linear"):
            raise ValueError(f"Invalid format: {format}. Use 'dense', 'sparse', or 'linear'.")
        logger.info(f"CONVERT HAMILTONIAN from {self.Ham_type} to {format}")
        self.Ham_type = format

    def get_sparsity(self):
        """Stored entries over N^2 (qmb_hamiltonian.py:614-619 counts the
        triplets before duplicates are summed; the device matrix keeps the
        summed ones)."""
        sparsity = self.nnz / (self.shape[0] ** 2)
        logger.info(f"SPARSITY: {round(sparsity, 16)}")
        return sparsity

    # -- spectral sums (qmb_hamiltonian.py:428-601)
    def full_eigenpairs(self):
        """(energies, eigenvector rows (N, N) on the device) of every level:
        the stored ones after diagonalize("full"), else one device solve over
        the whole sector (the full-basis Lanczos path up to 4096 states, the
        dense rocSOLVER factorization above), kept until the matrix changes.  The reference forms exp(-beta H) densely
        where the spectrum is not stored; every trace it takes is a sum over
        these levels."""
        import torch
        n = self.shape[0]
        if getattr(self, "n_eigs", 0) == n:
            return np.asarray(self.energies, dtype=float), torch.stack(
                [s.psi_device.to(torch.complex128) for s in self.Npsi])
        if getattr(self, "_all_pairs", None) is None:
            if n > 4096:
                evals, evecs = self.Ham.dense_eigh()
            else:
                from .device import random_uniform
                v0 = random_uniform(n, complex_values=self.Ham.is_complex, seed=0)
                evals, evecs, _, _ = self.Ham.lanczos(n, tol=1e-12, v0=v0)
            self._all_pairs = (np.asarray(evals, dtype=float), evecs.to(torch.complex128))
        return self._all_pairs

    def _levels(self):
        """The levels the reference's sums run over: the per-site energies when
        every eigenpair is stored (its full-spectrum branch, qmb_hamiltonian.py:
        446-447), otherwise the eigenvalues of H (its expm branch, :448-449:
        Tr f(H) = sum_i f(E_i))."""
        if getattr(self, "n_eigs", 0) == self.shape[0]:
            return np.asarray(self.Nenergies, dtype=float)
        return self.full_eigenpairs()[0]

    def partition_function(self, beta):
        """Z = sum_i exp(-beta e_i) (qmb_hamiltonian.py:428-452); ValueError
        unless Z > 0."""
        Z = float(np.sum(np.exp(-beta * self._levels())))
        if Z <= 0:
            raise ValueError(f"Z must be positive, not {Z}.")
        return Z

    def free_energy(self, beta):
        """-log(Z) / beta (qmb_hamiltonian.py:454-469)."""
        return -1 / beta * np.log(self.partition_function(beta))

    def thermal_average(self, beta):
        """sum_i e_i exp(-beta e_i) / Z (qmb_hamiltonian.py:471-489)."""
        e = self._levels()
        return float(e.dot(np.exp(-beta * e)) / self.partition_function(beta))

    def F_prime(self, beta):
        """-<e^2>_beta - <e>_beta^2, the slope the reference's Newton search
        uses (qmb_hamiltonian.py:491-508; kept as written there)."""
        e = self._levels()
        e2 = np.square(e).dot(np.exp(-beta * e)) / self.partition_function(beta)
        return float(-e2 - self.thermal_average(beta) ** 2)

    def get_beta(self, state, threshold=1e-10, max_iter=1000):
        """Newton search for beta with thermal_average(beta) = <state|H|state>
        (qmb_hamiltonian.py:510-564): from beta = 1e-7, beta <- beta - F/F',
        F = thermal_average(beta) - <H>, until the step is below threshold or
        after max_iter steps (a warning).  As in the reference, <H> is the
        total energy while the full-spectrum branch weighs per-site energies."""
        import torch
        st = state if hasattr(state, "data_ptr") else torch.as_tensor(
            np.asarray(state, dtype=np.complex128).reshape(-1), device="cuda")
        state_energy = QMB_state(st.to(torch.complex128), self.lvals, None).expectation_value(self.Ham)
        beta, accuracy, it = 1e-7, 1.0, 0
        while accuracy > threshold and it < max_iter:
            it += 1
            F = self.thermal_average(beta) - state_energy
            Fp = self.F_prime(beta)
            prev, beta = beta, beta - F / Fp
            accuracy = abs(beta - prev)
        if it >= max_iter:
            logger.warning("Maximum iterations reached without convergence.")
        logger.info(f"BETA {beta}")
        return beta

    def get_r_value(self):
        """Adjacent-gap ratios r_i = min(d_i, d_i+1) / max(d_i, d_i+1) of the
        sorted per-site spectrum; self.r_value = their mean over the middle
        80 % (qmb_hamiltonian.py:566-601).  Needs every eigenpair."""
        if getattr(self, "n_eigs", 0) != self.shape[0]:
            raise ValueError("The Hamiltonian has to be fully diagonalized.")
        d = np.diff(np.sort(np.asarray(self.Nenergies, dtype=float)))
        r = np.minimum(d[:-1], d[1:]) / np.maximum(d[:-1], d[1:])
        self.r_value = float(np.mean(r[int(0.1 * len(r)):int(0.9 * len(r))]))
        logger.info(f"R value {self.r_value}")
        return r

    def diagonalize(self, n_eigs, format, loc_dims, print_results=False, tol=1e-12,
                    max_basis=0, precision=L.PREC_FP64, v0=None, gauge=None):
        """Lowest n_eigs eigenpairs; Nenergies are energies per site
        (qmb_hamiltonian.py:260-332).  gauge = (real CSR of D^H H D, sector
        basis, per-site exponent table): solve that matrix and return the
        eigenvectors D x of H (QuantumModel.real_gauge_matrix)."""
        if isinstance(n_eigs, str):
            if n_eigs != "full":
                raise ValueError(f"n_eigs must be int or 'full', not {n_eigs}.")
            # every eigenpair: up to 4096 states the device solver's
            # full-basis path (the Krylov space is the whole sector, then one
            # host factorization of the N x N projected matrix); above, the
            # dense matrix factored on the device (rocSOLVER), as the
            # reference factors its dense array (qmb_hamiltonian.py:293-301)
            if self.shape[0] > 4096:
                return self._diagonalize_dense(loc_dims, print_results, gauge)
            n_eigs = int(self.shape[0])
        if not isinstance(n_eigs, (int, np.integer)) or n_eigs < 1:
            raise ValueError(f"n_eigs must be int or 'full', not {n_eigs}.")
        if n_eigs > self.shape[0]:
            raise ValueError(f"n_eigs must be < H.shape[0]={self.shape[0]}, got {n_eigs}")
        self.loc_dims = loc_dims
        self.n_eigs = int(n_eigs)
        H = self.Ham if gauge is None else gauge[0]
        cplx = H.is_complex
        if v0 is None:
            # default_rng(0).random(N), generated on the device (the solver
            # normalises its start block)
            from .device import random_uniform
            v0 = random_uniform(self.shape[0], complex_values=cplx, seed=0)
        elif gauge is not None:
            raise ValueError("a start vector is not taken by the real-gauge solve")
        else:
            v0 = np.asarray(v0, dtype=np.complex128 if cplx else np.float64)
            v0 = v0 / np.linalg.norm(v0)
        evals, evecs, resid, mvs = H.lanczos(self.n_eigs, tol=tol, max_basis=max_basis,
                                             precision=precision, v0=v0)
        if gauge is not None:
            evecs = gauge[1].quarter_phase(gauge[2], evecs)
        self.residuals = resid
        self.matvecs = mvs
        self._store_eigenpairs(evals, evecs, loc_dims, print_results)

    def time_evolution(self, initial_state, time_line, loc_dims, basis=None, d_max=None, m=30,
                       reorth=True, precision=L.PREC_FP64, store_states=True, gauge=None):
        """exp(-iHt) psi0 on a uniform time grid by device Krylov stepping
        (replaces expm_multiply in QMB_hamiltonian.time_evolution,
        qmb_hamiltonian.py:334-426).  psi_time[k] = exp(-i H t_k) psi0.  The
        stepping state (m+1 basis vectors) stays on the device for the next call
        with the same matrix and settings; release_krylov() frees it.
        gauge = (real CSR of D^H H D, sector basis, per-site exponent table):
        evolve D^H psi0 under that matrix (exp(-iHt) = D exp(-i D^H H D t) D^H)
        and map the states back through D; the site marginals are the same in
        both gauges."""
        import torch
        from .device import KrylovSession
        if self.Ham_type == "dense":
            return self._exact_time_evolution(initial_state, time_line, loc_dims)
        time_line = np.asarray(time_line, dtype=float)
        if time_line.size < 2:
            raise ValueError("time_line needs at least two points")
        dt = float(time_line[1] - time_line[0])
        if not np.allclose(np.diff(time_line), dt, rtol=1e-10, atol=1e-12):
            raise ValueError("time_line must be uniformly spaced")
        self.loc_dims = loc_dims
        N = self.shape[0]
        if hasattr(initial_state, "data_ptr"):
            psi = initial_state.to(device="cuda", dtype=torch.complex128).clone()
        else:
            psi = torch.as_tensor(np.asarray(initial_state, dtype=np.complex128), device="cuda").clone()
        if psi.numel() != N:
            raise ValueError("initial_state has the wrong length")
        H = self.Ham if gauge is None else gauge[0]
        if gauge is not None:
            gauge[1].quarter_phase(gauge[2], psi, direction=-1, out=psi)  # D^H psi0
        dm = int(d_max or 0) if basis is not None else 0
        key = (id(H), None if basis is None else id(basis), int(precision), int(m), bool(reorth), dm)
        ks = getattr(self, "_krylov", None)
        if ks is None or ks.key != key or ks.csr is not H:
            self.release_krylov()
            ks = self._krylov = KrylovSession(H, basis, precision, m, reorth, dm)
        if abs(time_line[0]) > 0:
            ks.run(psi, float(time_line[0]), 1)
        n_steps = time_line.size - 1
        hist = torch.empty((n_steps + 1, N), dtype=torch.complex128, device="cuda") if store_states else None
        marg = None
        if basis is not None:
            marg = np.zeros((n_steps + 1, basis.n_sites, d_max))
        err = np.zeros(max(n_steps, 1))
        ks.run(psi, dt, n_steps, marg, err, hist)
        if gauge is not None:  # back to the gauge of H
            gauge[1].quarter_phase(gauge[2], psi, direction=1, out=psi)
            if hist is not None:
                gauge[1].quarter_phase(gauge[2], hist, direction=1, out=hist)
        self.krylov_err = err[:n_steps]
        self.marginals = marg
        self.psi_final = psi
        self.psi_time = [QMB_state(hist[k], self.lvals, loc_dims) for k in range(n_steps + 1)] if store_states else None

    def _store_eigenpairs(self, evals, evecs, loc_dims, print_results):
        self.energies = evals
        nsites = prod(self.lvals)
        self.Nenergies = evals / nsites
        self.Npsi = [QMB_state(evecs[i], self.lvals, loc_dims) for i in range(self.n_eigs)]
        self.GSenergy = self.Nenergies[0]
        self.GSpsi = self.Npsi[0]
        if print_results:
            for e in evals:
                logger.info(f"{e:.10f}  {e / nsites:.10f}")

    def _diagonalize_dense(self, loc_dims, print_results, gauge):
        """n_eigs="full" above the Lanczos full-basis limit: the stored matrix
        (or the real-gauge one, eigenvectors mapped back) factored densely on
        the device."""
        self.loc_dims = loc_dims
        H = self.Ham if gauge is None else gauge[0]
        evals, evecs = H.dense_eigh()
        if gauge is not None:
            evecs = gauge[1].quarter_phase(gauge[2], evecs)
        self.n_eigs = int(self.shape[0])
        self.residuals = np.zeros(self.n_eigs)
        self.matvecs = np.zeros(2, dtype=np.int64)
        self._store_eigenpairs(evals, evecs, loc_dims, print_results)

    def _exact_time_evolution(self, initial_state, time_line, loc_dims):
        """The reference's path for ham_format="dense" (qmb_hamiltonian.py:
        363-376, exact_time_evolution :792-836): every eigenpair (solved
        first where not stored), then psi(t) = sum_i e^{-i E_i t} <E_i|psi0> E_i
        at each time of the grid (any spacing), as one product on the device."""
        import torch
        if getattr(self, "n_eigs", 0) != self.shape[0]:
            self.diagonalize("full", "dense", loc_dims)
        self.loc_dims = loc_dims
        en, E = self.full_eigenpairs()
        if hasattr(initial_state, "data_ptr"):
            psi0 = initial_state.to(device="cuda", dtype=torch.complex128).reshape(-1)
        else:
            psi0 = torch.as_tensor(np.asarray(initial_state, dtype=np.complex128).reshape(-1), device="cuda")
        if psi0.numel() != self.shape[0]:
            raise ValueError("initial_state has the wrong length")
        t = np.asarray(time_line, dtype=float).reshape(-1)
        coef = torch.mv(E.conj(), psi0)
        ph = torch.as_tensor(np.exp(-1j * np.outer(t, en)), device="cuda")
        hist = (ph * coef[None, :]) @ E
        self.krylov_err = np.zeros(max(t.size - 1, 1))
        self.marginals = None
        self.psi_final = hist[-1].clone()
        self.psi_time = [QMB_state(hist[k], self.lvals, loc_dims) for k in range(t.size)]

    def release_krylov(self):
        """Frees the device state kept by time_evolution."""
        ks = getattr(self, "_krylov", None)
        if ks is not None:
            ks.close()
        self._krylov = None


class QuantumModel:
    """Model base (edlgt/models/quantum_model.py:120-615)."""

    def __init__(self, lvals, has_obc, ham_format="sparse", basis_projector=None):
        if basis_projector is not None:
            raise NotImplementedError("basis_projector is not supported by the device engine")
        self.lvals = list(lvals)
        self.has_obc = list(has_obc)
        self.lattice = Lattice(self.lvals, self.has_obc)
        self.dim = len(self.lvals)
        self.directions = "xyz"[: self.dim]
        self.n_sites = prod(self.lvals)
        self.ham_format = ham_format
        self.momentum_basis = None
        # True: build the Hamiltonian as a matrix-free operator; False: always
        # a stored CSR; "auto": from the stored entries counted on the device,
        # a stored CSR while it takes at most half the free device memory, a
        # dictionary-coded one while the full build fits and the coded form
        # takes at most half, else matrix-free
        self.matrix_free = "auto"
        self.staggered_basis = False
        self.basis = None
        self.res = {}

    # -- per-site operator tables (quantum_model.py:400-502)
    def _project(self, site, op_builder, labels):
        self.lattice_labels = labels
        self.loc_dims = np.array([site.dim(l) for l in labels], dtype=int)
        d_max = int(self.loc_dims.max())
        per_label = {l: op_builder(l) for l in sorted(set(labels))}
        names = list(next(iter(per_label.values())).keys())
        self.ops = {}
        for name in names:
            arr = np.zeros((self.n_sites, d_max, d_max), dtype=np.complex128)
            for j, l in enumerate(labels):
                d = self.loc_dims[j]
                arr[j, :d, :d] = per_label[l][name]
            self.ops[name] = arr
        self._label_ops = per_label
        self.gauge_states_per_site = [site.values[l] for l in labels]

    @property
    def sector_configs(self):
        return None if self.basis is None else self.basis.configs()

    @property
    def sector_dim(self):
        return None if self.basis is None else self.basis.n_states

    def default_params(self):
        self.H = QMB_hamiltonian(self.lvals, size=self.sector_dim)
        self.H.psi_time = None
        self.H.marginals = None

    # -- Hamiltonian program: tables + placements
    def _new_program(self):
        self._tables = []
        self._table_ids = {}
        self._placements = []

    def _table(self, name, site):
        key = (name, self.lattice_labels[site])
        if key not in self._table_ids:
            self._table_ids[key] = len(self._tables)
            self._tables.append(self.ops[name][site])
        return self._table_ids[key]

    def _add_local(self, name, strength, sites=None):
        """LocalTerm.get_Hamiltonian (local_term.py:55-117): diagonal entries
        O[a][a] with |O[a][a]| >= 1e-10, times strength, on the masked sites."""
        for i in range(self.n_sites) if sites is None else sites:
            self._placements.append(dict(sites=[i], tables=[self._table(name, i)], strength=strength))

    def _add_nbody(self, names, site_lists, strength, add_dagger):
        """TwoBodyTerm / PlaquetteTerm / NBodyTerm (twobody_term.py:116-158,
        plaquette_term.py:116-163): one placement per site tuple, plus the
        Hermitian-conjugate rows when add_dagger."""
        for sl in site_lists:
            tabs = [self._table(n, s) for n, s in zip(names, sl)]
            self._placements.append(dict(sites=list(sl), tables=tabs, strength=strength, dagger=0))
            if add_dagger:
                self._placements.append(dict(sites=list(sl), tables=tabs, strength=strength, dagger=1))

    def _storage_mode(self, tables, vt, placements=None):
        placements = self._placements if placements is None else placements
        if self.momentum_basis is not None or self.matrix_free is False:
            return "stored"
        if self.matrix_free is True:
            return "free"
        import torch
        from .device import storage_plan, stored_entries, stored_entries_bound
        free, _ = torch.cuda.mem_get_info()
        bound = stored_entries_bound(tables, placements, self.sector_dim)
        if storage_plan(self.sector_dim, bound, vt, free) == "stored":
            return "stored"  # even the host bound fits: no device count needed
        nnz = stored_entries(self.basis, tables, placements)
        return storage_plan(self.sector_dim, nnz, vt, free)

    def _make_csr(self, tables, placements, vt):
        mode = self._storage_mode(tables, vt, placements)
        csr = CSRMatrix(self.basis, tables, placements, vt, matrix_free=mode == "free")
        if mode == "coded":
            # values replaced by dictionary codes (products bit for bit the
            # same); a matrix with too many distinct values is rebuilt
            # matrix-free instead
            if csr.compress() > 0:
                csr.release_values()
                logger.info("Hamiltonian stored dictionary-coded (values released)")
            else:
                del csr
                mode = "free"
                csr = CSRMatrix(self.basis, tables, placements, vt, matrix_free=True)
        if mode == "free":
            logger.info("Hamiltonian built matrix-free (entries generated in every product)")
        return csr

    def _keep_program(self):
        """Records the term program of the Hamiltonian being built (the
        real-gauge matrix is built from it later; build_multi_term rewrites the
        working program) and drops the real-gauge matrix of the previous build."""
        self._h_program = (list(self._tables), [dict(p) for p in self._placements])
        self._rg = None

    def _build_program(self, real):
        self._keep_program()
        tables = np.stack(self._tables)
        vt = L.VALUE_F64 if real else L.VALUE_C128
        csr = self._make_csr(tables, self._placements, vt)
        if self.momentum_basis is not None:
            # B^H H B in the momentum sector (the reference's momentum triplets,
            # translational_sym.py:1394-1969, behind QMB_hamiltonian.build)
            csr = csr.momentum(self.momentum_basis)
            self.H.shape = (csr.n, csr.n)
        self.H.set_matrix(csr, self.ham_format)
        return csr

    # -- momentum sectors (quantum_model.py:213-276)
    def set_momentum_sector(self, k_unit_cell_size, k_vals, TC_symmetry=False):
        """Restricts the following builds to momentum k_vals of the block
        translations by k_unit_cell_size sites (axis order and coordinates of
        the reference's translational_sym module).  The Hamiltonian becomes
        B^H H B on the n_cols momentum states; measure_observables maps the
        eigenvectors back through B."""
        self.momentum_basis = self._momentum_projector(k_unit_cell_size, k_vals, TC_symmetry)
        logger.info(f"Momentum basis shape ({self.momentum_basis.n_rows},{self.momentum_basis.n_cols})")

    def _momentum_projector(self, k_unit_cell_size, k_vals, TC_symmetry):
        from .device import MomentumBasis
        if self.basis is None:
            raise ValueError("symmetry sector_configs not defined yet")
        if self.has_obc[0]:
            raise ValueError("Momentum is not conserved with OBC.")
        k_vals = np.atleast_1d(np.asarray(k_vals, dtype=np.int32))
        if k_vals.size != self.dim:
            raise ValueError(f"expected k_vals shape {self.dim}, got {k_vals.shape}")
        uc = np.ascontiguousarray(k_unit_cell_size, dtype=np.int32)
        if TC_symmetry:
            tc = getattr(self, "tc_symmetry_tables", None)
            if tc is None or self.dim != 1:
                raise NotImplementedError("TC_symmetry momentum sectors exist for the 1D SU(2) chain only")
            return MomentumBasis(self.basis, self.lvals, uc, k_vals, tc_map=tc[0], tc_sign=tc[1])
        return MomentumBasis(self.basis, self.lvals, uc, k_vals)

    def set_momentum_pair(self, k_left, k_right, k_unit_cell_size, TC_symmetry):
        """Two momentum projectors P_kL, P_kR for rectangular blocks
        P_kL^dagger O P_kR of operators built afterwards
        (quantum_model.py:278-348): kept in self.momentum_pair; the sector
        basis and its Hamiltonian are unchanged."""
        mb_l = self._momentum_projector(k_unit_cell_size, k_left, TC_symmetry)
        mb_r = self._momentum_projector(k_unit_cell_size, k_right, TC_symmetry)
        logger.info(f"Combined Momenta {k_left} {k_right}")
        self.momentum_pair = {"L": mb_l, "R": mb_r, "k_left": np.atleast_1d(np.asarray(k_left, dtype=np.int32)),
                              "k_right": np.atleast_1d(np.asarray(k_right, dtype=np.int32))}

    def check_momentum_pair(self):
        """||P_L^H P_L - I||_F and ||P_R^H P_R - I||_F <= 1e-12, and, for
        different momenta, ||P_L^H P_R||_F <= 1e-12 (quantum_model.py:350-398);
        ValueError otherwise."""
        import scipy.sparse as sps
        from scipy.sparse.linalg import norm
        mp = getattr(self, "momentum_pair", None)
        if mp is None:
            raise ValueError("no momentum pair set (set_momentum_pair)")
        PL, PR = mp["L"].to_scipy().tocsc(), mp["R"].to_scipy().tocsr()
        nl = norm(PL.conj().T @ PL - sps.identity(PL.shape[1]))
        nr = norm(PR.conj().T @ PR - sps.identity(PR.shape[1]))
        nmix = norm(PL.conj().T @ PR)
        logger.info(f"norm: (PL^dag @ PL) -1: {nl}")
        logger.info(f"norm: (PR^dag @ PR) -1: {nr}")
        logger.info(f"norm: (PL^dag @ PR): {nmix}")
        if nl > 1e-12:
            raise ValueError("PL is not a projector")
        if nr > 1e-12:
            raise ValueError("RL is not a projector")
        if not np.array_equal(mp["k_left"], mp["k_right"]) and nmix > 1e-12:
            raise ValueError("The two basis are not orthogonal")

    def get_abelian_symmetry_sector(self, global_ops=None, global_sectors=None, global_sym_type="U",
                                    link_ops=None, link_sectors=None, nbody_ops=None, nbody_sectors=None,
                                    nbody_sites_list=None, nbody_sym_type=None):
        """The symmetry sector of abelian generators given as per-site operator
        tables (quantum_model.py:504-604): global generators with their
        sectors ("U" sums or "Z" products), link generators [O_plus, O_minus]
        per direction on the lattice's link pairs, and n-body generators on
        site lists (used with link generators only, as in the reference).  The
        diagonals of the tables define the constraints; the sector is
        enumerated on the device and replaces this model's basis (the
        Hamiltonian and the per-site tables stay as they are).  Without any
        generator the sector is the whole product space."""
        def diag(op):
            return np.real(np.einsum("sii->si", np.asarray(op)))

        glob = link = nbody = None
        if global_ops is not None:
            glob = (np.stack([diag(o) for o in global_ops]), np.asarray(global_sectors, dtype=float),
                    global_sym_type)
        if link_ops is not None:
            ld = np.stack([np.stack([diag(pair[0]), diag(pair[1])]) for pair in link_ops])
            link = (ld, np.asarray(link_sectors, dtype=float), self.lattice.link_pairs())
            if nbody_ops is not None and global_ops is None:
                nbody = (np.stack([diag(o) for o in nbody_ops]), np.asarray(nbody_sectors, dtype=float),
                         [list(map(int, sl)) for sl in nbody_sites_list], nbody_sym_type or "U")
        self.basis = SectorBasis(self.loc_dims, glob=glob, link=link, nbody=nbody)
        if self.basis.n_states == 0:
            raise ValueError("No configurations found for the given symmetry sectors")
        logger.info(f"SEC DIM: {self.basis.n_states}")
        self.momentum_basis = None
        self.default_params()
        return self.basis

    def _basis_Pk_as_csr(self):
        """The momentum projector B (sector rows x momentum columns) as SciPy
        CSR, or None without a momentum sector (quantum_model.py:626-632)."""
        return None if self.momentum_basis is None else self.momentum_basis.to_scipy().tocsr()

    def _basis_Pk_as_csc(self):
        """The same projector as SciPy CSC (quantum_model.py:615-624)."""
        return None if self.momentum_basis is None else self.momentum_basis.to_scipy().tocsc()

    def _project_state_with_basis(self, state_realspace):
        """B^H x: sector-basis coefficients to the momentum basis
        (quantum_model.py:634-671), on the device; the identity (as complex128)
        without a momentum sector.  Returns a host array."""
        import torch
        x = np.asarray(state_realspace.cpu().numpy() if hasattr(state_realspace, "data_ptr") else state_realspace)
        if self.momentum_basis is None:
            return x.astype(np.complex128, copy=False)
        t = torch.as_tensor(np.ascontiguousarray(x, dtype=np.complex128).reshape(-1), device="cuda")
        return self.momentum_basis.project(t).cpu().numpy()

    def momentum_basis_projection(self, operator):
        """B^H A B in the momentum basis (quantum_model.py:673-713).  "H": the
        device Hamiltonian, when it was built in the sector basis before the
        momentum sector was set, is projected on the device in place (built
        after set_momentum_sector it already is B^H H B); an operator given as a
        dense or SciPy matrix over the sector basis is projected with SciPy, as
        the reference does, and returned as CSR."""
        import scipy.sparse as sps
        mb = self.momentum_basis
        if mb is None:
            raise ValueError("Basis projector B is not set.")
        if isinstance(operator, str) and operator == "H":
            H = self.H.Ham
            if H.n == mb.n_rows and getattr(H, "momentum_basis", None) is None:
                P = H.momentum(mb)
                self.H.shape = (P.n, P.n)
                self.H.set_matrix(P, self.H.Ham_type)
            elif H.n != mb.n_cols:
                raise ValueError(f"H has {H.n} rows, the momentum basis maps {mb.n_rows} -> {mb.n_cols}")
            return None
        A = sps.csr_matrix(operator)
        B = self._basis_Pk_as_csc()
        return (B.conj().T @ (A @ B)).tocsr()

    def _sector_state(self, psi_device):
        """A state over the sector basis (momentum states expanded through B)."""
        mb = self.momentum_basis
        if mb is not None and psi_device.numel() == mb.n_cols and mb.n_cols != mb.n_rows:
            return mb.expand(psi_device)
        return psi_device

    def diagonalize_Hamiltonian(self, n_eigs, format, print_results=False, real_gauge=False, **kw):
        """quantum_model.py:606-611.  real_gauge=True solves D^H H D instead of H
        when a diagonal phase D makes it real (real_gauge_matrix): the same
        eigenvalues with real float64 vectors and values (about half the bytes
        per product and per orthogonalisation), the eigenvectors mapped back
        through D on the device.  Where no such phase applies the complex
        matrix is solved; H.real_gauge tells which ran."""
        gauge = None
        if real_gauge:
            rg = self.real_gauge_matrix()
            if rg is not None:
                gauge = (rg[0], self.basis, rg[1])
        self.H.diagonalize(n_eigs, format, self.loc_dims, print_results, gauge=gauge, **kw)
        self.H.real_gauge = gauge is not None
        self.n_eigs = self.H.n_eigs
        self.res["energy"] = self.H.Nenergies

    # -- real gauge (no counterpart in the reference: its complex solve gives
    # the same eigenpairs)
    def _gauge_weights(self):
        """(a, name): per-site weights a_j and the number operator n_j of a
        diagonal phase D = diag(i^(sum_j a_j n_j)) under which the model's
        Hamiltonian is real, or None when the model knows no such phase."""
        return None

    def _real_gauge_placements(self, program=None):
        """The program of D^H H D (of the last build_Hamiltonian, or the given
        (tables, placements)): each placement's strength times
        i^(-sum_f a_f q_f), where q_f is the change of n (bra minus ket) that
        factor f's table makes at its site.  A placement maps a ket c to a bra r
        with s_r - s_c = sum_f a_f q_f (s = sum_j a_j n_j), so every entry of
        the placement takes the same factor i^(s_c - s_r); its Hermitian
        conjugate takes the conjugate factor, which the assembly's conj(strength)
        of dagger placements already gives.  None when a table changes n by
        different amounts on different entries (no fixed factor)."""
        w = self._gauge_weights()
        prog = getattr(self, "_h_program", None) if program is None else program
        if w is None or prog is None or self.momentum_basis is not None:
            return None
        tables, placements = prog
        a, name = w
        nd = np.real(np.einsum("sii->si", self.ops[name]))
        quarter = (1.0, 1j, -1.0, -1j)
        charge = {}
        out = []
        for P in placements:
            e = 0
            for site, t in zip(P["sites"], P["tables"]):
                if (t, site) not in charge:
                    bra, ket = np.nonzero(np.abs(tables[t]) > 1e-10)
                    q = nd[site][bra] - nd[site][ket]
                    fixed = q.size == 0 or (np.all(np.abs(q - q[0]) <= 1e-9) and abs(q[0] - round(q[0])) <= 1e-9)
                    charge[(t, site)] = (int(round(q[0])) if q.size else 0) if fixed else None
                if charge[(t, site)] is None:
                    return None
                e += int(a[site]) * charge[(t, site)]
            Q = dict(P)
            Q["strength"] = P["strength"] * quarter[(-e) % 4]
            out.append(Q)
        return out

    def real_gauge_matrix(self):
        """(float64 CSR of D^H H D, per-site exponent table a_j n_j(state))
        for the last build_Hamiltonian, or None when the model has no such
        phase, a momentum sector is set, or D^H H D is not real (QED with matter
        on a lattice periodic along x or z with an odd extent).  Built once per
        build_Hamiltonian by the same assembly, which rejects any entry with an
        imaginary part above 1e-12."""
        cached = getattr(self, "_rg", None)
        if cached is not None:
            return cached or None
        self._rg = False
        placements = self._real_gauge_placements()
        if placements is None:
            return None
        try:
            csr = self._make_csr(np.stack(self._h_program[0]), placements, L.VALUE_F64)
        except ValueError:
            logger.info("no real gauge: D^H H D has complex entries")
            return None
        a, name = self._gauge_weights()
        nd = np.rint(np.real(np.einsum("sii->si", self.ops[name]))).astype(np.int64)
        site_exp = (np.asarray(a, dtype=np.int64)[:, None] * nd).astype(np.int32)
        self._rg = (csr, site_exp)
        return self._rg

    def time_evolution_Hamiltonian(self, initial_state, time_line, real_gauge=False, **kw):
        """quantum_model.py:613-615; per-step site marginals are kept in
        self.H.marginals so that dynamics observables need no extra pass.
        real_gauge=True evolves under the real matrix D^H H D when a diagonal
        phase D makes it real (real_gauge_matrix; complex states, float64
        values: 12 instead of 20 bytes per stored entry in every product) and
        maps the states back; H.real_gauge tells which ran."""
        d_max = int(self.loc_dims.max())
        # per-step marginals need sector-basis states: not in a momentum sector
        basis = self.basis if self.momentum_basis is None else None
        gauge = None
        if real_gauge:
            rg = self.real_gauge_matrix()
            if rg is not None:
                gauge = (rg[0], self.basis, rg[1])
        if self.H.Ham_type == "dense":
            gauge = None  # exact evolution from the stored spectrum
        self.H.time_evolution(initial_state, time_line, self.loc_dims, basis=basis, d_max=d_max, gauge=gauge, **kw)
        self.H.real_gauge = gauge is not None

    # -- observables (quantum_model.py:1158-1333)
    def get_observables(self, local_obs=[], twobody_obs=[], plaquette_obs=[], nbody_obs=[],
                        nbody_dist=[], twobody_axes=None):
        """quantum_model.py:1158-1235.  Local (diagonal) observables come from
        site marginals; two-body correlators, plaquettes and n-body terms from
        one device expectation pass each (edlgt_expval_terms)."""
        self.local_obs = list(local_obs)
        self.twobody_obs = [list(p) for p in twobody_obs]
        self.twobody_axes = twobody_axes
        self.plaquette_obs = [list(p) for p in plaquette_obs]
        self.nbody_obs = [list(p) for p in nbody_obs]
        if len(nbody_dist) < len(self.nbody_obs):
            raise ValueError("nbody_dist needs one distance list per n-body observable")
        self.nbody_dist = [list(d) for d in nbody_dist]
        for names, dist in zip(self.nbody_obs, self.nbody_dist):
            if len(dist) != len(names) - 1:
                raise ValueError(f"The distances length should be len(op_list)-1, not {len(dist)}")
            if any(len(d) != self.dim for d in dist):
                raise ValueError(f"Each entry in distances must have length len(lvals)={self.dim}.")
        if twobody_axes is not None:
            for ax in twobody_axes:
                if ax not in self.directions:
                    raise ValueError(f"axis should be in {self.directions}: got {ax}")
        self._obs_groups = {}
        for names in self.twobody_obs:
            if len(names) != 2:
                raise ValueError("a two-body observable names two operators")
            pairs = [(i, j) for i in range(self.n_sites) for j in range(i + 1, self.n_sites)]
            self._obs_groups["_".join(names)] = ("twobody", names, pairs)
        for names in self.plaquette_obs:
            axes = (names[-1][3], names[-1][6])
            self._obs_groups["_".join(names)] = ("plaquette", names, self.lattice.plaquettes(axes))
        for names, dist in zip(self.nbody_obs, self.nbody_dist):
            sl = [self.lattice.nbody_sites(i, dist) for i in range(self.n_sites)]
            self._obs_groups["_".join(names)] = ("nbody", names, [x for x in sl if x is not None])

    def _group_expvals(self, psi_device):
        """Complex <psi|O_p|psi> of every placement of every correlator group, in
        one device pass."""
        from .device import expval_terms
        tables, ids, placements, spans = [], {}, [], {}

        def tab(name, site):
            key = (name, self.lattice_labels[site])
            if key not in ids:
                ids[key] = len(tables)
                tables.append(self.ops[name][site])
            return ids[key]

        for key, (kind, names, site_lists) in self._obs_groups.items():
            a = len(placements)
            for sl in site_lists:
                placements.append(dict(sites=list(sl), tables=[tab(n, q) for n, q in zip(names, sl)]))
            spans[key] = (a, len(placements))
        if not placements:
            return {}
        vals = expval_terms(self.basis, np.stack(tables), placements, psi_device)
        return {k: vals[a:b] for k, (a, b) in spans.items()}

    def _marginals(self, psi_device):
        d_max = int(self.loc_dims.max())
        out = np.zeros((self.n_sites, d_max))
        vt = L.VALUE_C128 if psi_device.is_complex() else L.VALUE_F64
        L.check(L.load().edlgt_basis_marginals(self.basis.handle, L.ptr(psi_device.contiguous()), vt,
                                               d_max, out.ctypes.data, L.stream_handle()))
        return out

    def local_expvals(self, marg, name):
        """<O_s> per site of a diagonal local operator from site marginals
        (LocalTerm in the sector is diagonal-only: sym_qmb_operations.py:565-602)."""
        diag = np.real(np.einsum("sii->si", self.ops[name]))
        return np.einsum("sa,sa->s", marg, diag[:, : marg.shape[1]])

    def measure_observables(self, index, dynamics=False):
        import torch
        if not dynamics:
            state = self.H.Npsi[index]
            psi_dev = self._sector_state(state.psi_device)
            marg = self._marginals(psi_dev)
        else:
            state = self.H.psi_time[index] if self.H.psi_time is not None else None
            psi_dev = None if state is None else self._sector_state(state.psi_device)
            marg = self.H.marginals[index] if self.H.marginals is not None else self._marginals(psi_dev)
        for obs in self.local_obs:
            self.res[obs] = self.local_expvals(marg, obs)
        groups = getattr(self, "_obs_groups", {})
        if groups:
            vals = self._group_expvals(psi_dev)
            for key, (kind, names, site_lists) in groups.items():
                v = np.real(vals[key])
                if kind == "twobody":
                    # lattice_twobody_exp_val (twobody_term.py:232-264): symmetric, zero diagonal
                    corr = np.zeros((self.n_sites, self.n_sites))
                    for (i, j), x in zip(site_lists, v):
                        corr[i, j] = corr[j, i] = x
                    self.res[key] = corr
                elif kind == "plaquette":
                    # PlaquetteTerm.get_expval (plaquette_term.py:246-261): mean over plaquettes
                    self.res[key] = float(np.mean(v)) if v.size else float("nan")
                    self.res[key + "_obs"] = v
                else:
                    # NBodyTerm.get_expval (nbody_term.py:194-220): one value per valid start site
                    self.res[key] = v

    def link_avg(self, obs_name):
        """Average of res[f'{obs_name}_p{d}'] over the existing links (quantum_model.py:1271-1299)."""
        tot, cnt = 0.0, 0
        for i in range(self.n_sites):
            for d in self.directions:
                if self.lattice.step(i, d) is not None:
                    tot += self.res[f"{obs_name}_p{d}"][i]
                    cnt += 1
        return tot / cnt

    def stag_avg(self, obs_name, staggered_avg=None):
        """Mean of res[obs_name] over all / even / odd sites (quantum_model.py:1301-1333)."""
        v = np.asarray(self.res[obs_name])
        if staggered_avg is None:
            return float(np.mean(v))
        even = self.lattice.even_mask()
        return float(np.mean(v[even if staggered_avg == "even" else ~even]))

    def _get_partition(self, keep_indices):
        """Subsystem / environment configuration maps of the sector basis
        (quantum_model.py:715-791), cached by the kept site set."""
        if not hasattr(self, "_partition_cache"):
            self._partition_cache = {}
        key = tuple(sorted(keep_indices))
        if key not in self._partition_cache:
            cfg = self.sector_configs
            env = [i for i in range(self.n_sites) if i not in key]
            sub_u, sub_map = np.unique(cfg[:, list(key)], axis=0, return_inverse=True)
            env_u, env_map = np.unique(cfg[:, env], axis=0, return_inverse=True)
            sub_map, env_map = sub_map.ravel(), env_map.ravel()
            self._partition_cache[key] = dict(subsys_indices=list(key), env_indices=env,
                                              unique_subsys_configs=sub_u, unique_env_configs=env_u,
                                              subsys_map=sub_map, env_map=env_map,
                                              subsys_dim=int(sub_map.max()) + 1, env_dim=int(env_map.max()) + 1)
        return self._partition_cache[key]

    def build_projector_from_sector_to_fullspace(self, indices):
        """Column isometry from the subsystem's sector configurations to the
        subsystem's full product space (quantum_model.py:793-824): entry 1 at
        (mixed-radix index of configuration j, rightmost site fastest; j)."""
        import scipy.sparse as sps
        u = self._get_partition(indices)["unique_subsys_configs"].astype(np.int64)
        dims = np.asarray(self.loc_dims, dtype=np.int64)[sorted(indices)]
        strides = np.ones(dims.size, dtype=np.int64)
        for k in range(dims.size - 2, -1, -1):
            strides[k] = strides[k + 1] * dims[k + 1]
        rows = u @ strides
        n = u.shape[0]
        return sps.csc_matrix((np.ones(n), (rows, np.arange(n))), shape=(int(np.prod(dims)), n))

    def measure_fidelity(self, state, index, dynamics=False, print_value=False):
        """|<state|psi_index>|^2 with an eigenstate or a time-evolved state
        (quantum_model.py:864-902); state: host array or device tensor over
        the Hamiltonian's basis."""
        import torch
        ref = self.H.Npsi[index] if not dynamics else self.H.psi_time[index]
        n = self.H.shape[0]
        if hasattr(state, "data_ptr"):
            st = state.to(device="cuda", dtype=torch.complex128).reshape(-1)
        else:
            st = torch.as_tensor(np.asarray(state, dtype=np.complex128).reshape(-1), device="cuda")
        if st.numel() != n:
            raise ValueError(f"len(state) must be {n} not {st.numel()}")
        fid = float(abs(torch.vdot(st, ref.psi_device.to(torch.complex128)).item()) ** 2)
        if print_value:
            logger.info("----------------------------------------------------")
            logger.info(f"FIDELITY: {fid}")
        return fid

    # -- ensemble averages over a fully diagonalised sector (quantum_model.py:904-1156)
    def _density_values(self, rho, local_obs_list, special_norms, staggered_avgs, prefix):
        """The reference's per-observable reductions (mean over sites,
        staggered mean, or dot with a normalisation vector / L) of the
        diagonal local observables in the density rho(r) (real device tensor),
        from one site-marginal pass over sqrt(rho).  Both reductions are linear
        in the per-site values, so a weighted sum of eigenstate densities gives
        the weighted sum of the eigenstates' values."""
        import torch
        marg = self._marginals(torch.sqrt(torch.clamp(rho, min=0.0)).contiguous())
        out = {}
        for obs in local_obs_list:
            v = self.local_expvals(marg, obs)
            if special_norms.get(obs) is not None:
                val = float(np.dot(v, special_norms[obs]) / self.n_sites)
            elif staggered_avgs.get(obs) is not None:
                self.res[obs] = v
                val = self.stag_avg(obs, staggered_avgs[obs])
            else:
                val = float(np.mean(v))
            out[f"{prefix}_{obs}"] = val
        return out

    def _eigvec_matrix(self):
        import torch
        return torch.stack([s.psi_device.to(torch.complex128) for s in self.H.Npsi])

    def _ensemble_rho(self, E, members, coef, bstart):
        """rho(r) = sum_b |sum_{t in block b} coef_t S E[members_t](r)|^2 on the
        device (edlgt_ensemble_density), S the map to the sector basis (the
        identity, or B for a momentum basis, whose member rows are expanded
        first)."""
        import torch
        from .device import ensemble_density
        members = np.asarray(members, dtype=np.int64)
        if self._sector_state(E[0]) is not E[0]:
            E = torch.stack([self._sector_state(E[int(m)].contiguous()) for m in members])
            members = np.arange(members.size, dtype=np.int64)
        return ensemble_density(E, members, coef, bstart)

    def _state_tensor(self, state):
        import torch
        if hasattr(state, "data_ptr"):
            st = state.to(device="cuda", dtype=torch.complex128).reshape(-1).contiguous()
        else:
            st = torch.as_tensor(np.asarray(state, dtype=np.complex128).reshape(-1), device="cuda")
        if st.numel() != self.H.shape[0]:
            raise ValueError(f"len(state) must be {self.H.shape[0]} not {st.numel()}")
        return st

    def microcanonical_avg(self, local_obs_list, state, special_norms=None, staggered_avgs=None):
        """Microcanonical averages of diagonal local observables
        (quantum_model.py:942-1045): the shell |e_i - <H>/L| < sqrt(<H^2> -
        <H>^2)/L of the per-site eigenvalues around the state, equal weights.
        Returns (coherent shell state (host), {"ME_<obs>": value}); <H> and
        <H^2> = |H psi|^2 from one device product."""
        import torch
        special_norms = special_norms or {}
        staggered_avgs = staggered_avgs or {}
        st = self._state_tensor(state)
        Hs = self.H.Ham.matvec(st) if self.H.Ham.is_complex else torch.complex(
            self.H.Ham.matvec(st.real.contiguous()), self.H.Ham.matvec(st.imag.contiguous()))
        Eq = float(torch.vdot(st, Hs).real.item())
        E2q = float(torch.vdot(Hs, Hs).real.item())
        e_q, d_e = Eq / self.n_sites, np.sqrt(max(E2q - Eq ** 2, 0.0)) / self.n_sites
        shell = np.nonzero(np.abs(np.asarray(self.H.Nenergies) - e_q) < d_e)[0]
        if shell.size == 0:
            raise ValueError(f"ME shell empty: no eigstate |e_i-e_q|<delta_e={d_e}.")
        E = self._eigvec_matrix()
        psi_th = (E[torch.as_tensor(shell, device="cuda")].sum(dim=0) / np.sqrt(shell.size)).cpu().numpy()
        rho = self._ensemble_rho(E, shell, np.full(shell.size, 1.0 / np.sqrt(shell.size), dtype=complex),
                                 np.arange(shell.size + 1))
        return psi_th, self._density_values(rho, local_obs_list, special_norms, staggered_avgs, "ME")

    def diagonal_avg(self, local_obs_list, state, special_norms=None, staggered_avgs=None, tol_deg=1e-10):
        """Diagonal-ensemble averages (quantum_model.py:1047-1156): weights
        |<E_i|psi>|^2, or, where eigenvalues are degenerate within tol_deg, the
        projections Psi_b of the state on each degenerate block (sum_b
        <Psi_b|O|Psi_b>).  Needs every eigenpair."""
        import torch
        if self.n_eigs != self.H.shape[0]:
            raise ValueError(f"Need all H eigvals {self.H.shape[0]}, not only {self.n_eigs}")
        special_norms = special_norms or {}
        staggered_avgs = staggered_avgs or {}
        st = self._state_tensor(state)
        E = self._eigvec_matrix()
        coef = torch.mv(E.conj(), st).cpu().numpy()
        block, n_b = build_energy_block_ids(self.H.Nenergies, tol_deg)
        # levels are contiguous in the ascending spectrum: block b = [starts[b], starts[b+1])
        starts = np.concatenate([[0], np.nonzero(np.diff(block))[0] + 1, [block.size]])
        rho = self._ensemble_rho(E, np.arange(block.size), coef, starts)
        return self._density_values(rho, local_obs_list, special_norms, staggered_avgs, "DE")

    def canonical_avg(self, local_obs, beta):
        """Tr(O e^{-beta H}) / (Z L) with O the sum over sites of a diagonal local
        operator (quantum_model.py:912-940), as a sum over every level with
        weights e^{-beta (E_i - E_0)} (the shift cancels in the ratio): the
        stored eigenpairs after diagonalize_Hamiltonian("full", ...), else one
        device solve of the whole sector (QMB_hamiltonian.full_eigenpairs) where
        the reference forms e^{-beta H} densely."""
        import torch
        en, E = self.H.full_eigenpairs()
        w = np.exp(-beta * (en - en.min()))
        rho = self._ensemble_rho(E, np.arange(en.size), np.sqrt(w / w.sum()).astype(complex), np.arange(en.size + 1))
        vals = self._density_values(rho, [local_obs], {}, {}, "C")
        return vals[f"C_{local_obs}"]

    def compute_expH(self, beta):
        """exp(-beta H) as a SciPy CSC matrix (quantum_model.py:904-906), from
        every level: V^T diag(e^{-beta E}) conj(V) on the device."""
        import scipy.sparse as sps
        import torch
        en, E = self.H.full_eigenpairs()
        w = torch.as_tensor(np.exp(-beta * en), device="cuda", dtype=torch.complex128)
        M = (E.transpose(0, 1) * w[None, :]).contiguous() @ E.conj().contiguous()
        return sps.csc_matrix(M.cpu().numpy())

    def get_thermal_beta(self, state, threshold):
        """Effective inverse temperature of a state (quantum_model.py:908-910)."""
        return self.H.get_beta(state, threshold)

    # -- link symmetries of measured fields (local_term.py:240-284)
    def check_link_symmetry(self, axis, obs1, obs2, value=0, sign=1):
        """res[obs1][i] + sign * res[obs2][j] == value (within 1e-10) on every
        bond (i, j) along axis, from the measured per-site values
        (local_term.py:240-284); ValueError naming the first violating site."""
        a, b = np.asarray(self.res[obs1]), np.asarray(self.res[obs2])
        for i in range(self.n_sites):
            j = self.lattice.step(i, axis)
            if j is None:
                continue
            if abs(a[i] + sign * b[j] - value) > 1e-10:
                logger.info(f"{a[i]}")
                logger.info(f"{b[j]}")
                raise ValueError(f"{axis}-Link Symmetry is violated at index {i}")
        logger.info("----------------------------------------------------")
        logger.info(f"{axis}-LINK SYMMETRY IS SATISFIED")

    # -- numeric mode (quantum_model.py:1342-1409)
    @staticmethod
    def _normalize_dtype_mode(dtype_mode, allow_auto=True):
        """bool (True = complex), "real" / "complex" / "auto" (any case) or
        None (= "auto") to the mode's label; TypeError / ValueError otherwise."""
        if isinstance(dtype_mode, (bool, np.bool_)):
            return "complex" if bool(dtype_mode) else "real"
        if dtype_mode is None:
            if allow_auto:
                return "auto"
            raise ValueError("dtype_mode cannot be None in this context")
        if not isinstance(dtype_mode, str):
            raise TypeError(f"dtype_mode must be bool/str/None, got {type(dtype_mode)}")
        mode = dtype_mode.strip().lower()
        valid = {"real", "complex", "auto"} if allow_auto else {"real", "complex"}
        if mode not in valid:
            raise ValueError(f"dtype_mode must be one of {sorted(valid)}, got {dtype_mode!r}")
        return mode

    def configure_dtype_mode(self, dtype_mode="auto", auto_mode="complex"):
        """The numeric mode ("real" or "complex") that dtype_mode selects,
        auto_mode standing in for "auto".  The device build takes its mode
        per call (build_Hamiltonian(..., dtype_mode=...): float64 values only
        where every entry is real), so this records the choice in
        self.dtype_mode and returns it."""
        mode = self._normalize_dtype_mode(dtype_mode, allow_auto=True)
        if mode == "auto":
            mode = self._normalize_dtype_mode(auto_mode, allow_auto=False)
        self.dtype_mode = mode
        return mode

    # -- configuration printout (QED_model.py:901-946, SU2_model.py:831-871)
    _pad_state_values = True  # QED pads each local value to the column width

    def _local_state_labels(self):
        """Columns of a local basis state: bg (with background charges), M
        (with matter), then the links -x.. and +x.. of the lattice's axes."""
        labels = ["bg"] if getattr(self, "background", 0) > 0 else []
        if not getattr(self, "pure_theory", True):
            labels.append("M")
        dirs = "xyz"[: int(self.dim)]
        return labels + [f"-{d}" for d in dirs] + [f"+{d}" for d in dirs]

    def _format_header(self, entry_width):
        return "[" + " ".join(f"{lab:>{entry_width}s}" for lab in self._local_state_labels()) + " ]"

    def print_state_config(self, config, amplitude=None):
        """Logs a configuration site by site: the local state index and its
        gauge-invariant basis state (matter, link values)."""
        if getattr(self, "spin", None) == "integrated":
            raise NotImplementedError("It does not work with spin='integrated'")
        logger.info("----------------------------------------------------")
        msg = f"CONFIG {config}"
        if amplitude is not None:
            msg += f" |psi|^2={np.abs(amplitude) ** 2:.8f}"
        logger.info(msg)
        logger.info("----------------------------------------------------")
        max_abs = max(1, int(abs(getattr(self, "spin", 1))), int(abs(getattr(self, "background", 0))))
        width = len(str(max_abs)) + 2
        logger.info(f"{'':>19s}{self._format_header(width)}")
        for ii, idx in enumerate(config):
            st = self.gauge_states_per_site[ii][int(idx)]
            if self._pad_state_values:
                body = " ".join(f"{int(round(v)):>{width}d}" for v in st)
            else:
                body = " ".join(f"{v}" for v in st)
            logger.info(f"SITE {ii:>2d} state {int(idx):>3d}: [{body} ]")

    def get_qmb_state_from_configs(self, configs):
        """Equal-weight normalised superposition of basis configurations
        (quantum_model.py:826-862); raises ValueError for a config outside
        the sector."""
        import torch
        idx = self.basis.lookup(np.asarray(configs))
        state = torch.zeros(self.sector_dim, dtype=torch.complex128, device="cuda")
        state[torch.as_tensor(idx, device="cuda")] = complex(1 / np.sqrt(len(idx)), 0)
        return state


def _integrated_operators():
    """Pauli matrices in the reference's normalisation (spin_operators.py:62-72:
    Sz = diag(1, -1), Sp = |0><1|) with N = (Sz + I)/2 and N_zero = I - N
    (QED_operators.py:541-545)."""
    I = np.eye(2)
    Sz = np.diag([1.0, -1.0])
    Sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    Sm = Sp.T.copy()
    N = 0.5 * (Sz + I)
    return {"I": I, "Sz": Sz, "Sp": Sp, "Sm": Sm, "Sx": Sp + Sm, "Sy": -1j * (Sp - Sm), "N": N, "N_zero": I - N}


class _PauliSite:
    """Two-state site of integrated QED (no gauge states)."""
    values = {"site": np.arange(2)[:, None]}

    def dim(self, label):
        return 2


class QED_Model(QuantumModel):
    """U(1) lattice gauge theory with optional staggered fermions
    (edlgt/models/QED_model.py:20-394, 719-776)."""

    def __init__(self, spin, pure_theory, bg_list=None, plaq_basis=False, link_symmetries=True,
                 get_only_bulk=False, **kwargs):
        super().__init__(**kwargs)
        # static background charges (QED_model.py:56-59, 96-101): one integer
        # charge per site; all zero is the same as none
        self.bg_list = None
        self.background = 0
        bg = None
        if bg_list is not None:
            bg = np.asarray(bg_list, dtype=float)
            if bg.shape != (self.n_sites,):
                raise ValueError(f"len(bg_list) must match n_sites = {self.n_sites}: got {bg.shape}")
        if isinstance(spin, str):
            # gauge fields integrated out by Gauss's law (QED_model.py:64-82)
            if spin != "integrated":
                raise ValueError(f"spin must be integer or 'integrated': got {spin}")
            if pure_theory or self.dim != 1:
                raise ValueError("QED with integrated gauge fields is only in 1D with matter")
            self.bg_list = None if bg is None else bg.tolist()
            self._init_integrated()
            return
        if bg is not None:
            self.background = int(np.max(np.abs(bg)))
            if self.background != 0:
                self.bg_list = bg.tolist()
        if plaq_basis:
            raise NotImplementedError("the plaquette basis of the pure 2D theory is not built; "
                                      "use the dressed-site basis")
        if get_only_bulk and any(self.has_obc):
            # only the bulk site bases exist: border sites have none to project on
            raise ValueError("get_only_bulk=True keeps no border bases: it needs periodic boundaries")
        if not isinstance(spin, (int, np.integer)) or spin < 1:
            raise NotImplementedError("spin must be a positive integer truncation")
        self.spin = int(spin)
        self.pure_theory = bool(pure_theory)
        self.staggered_basis = not self.pure_theory
        # link_symmetries=False: the sector drops the link constraints and H
        # carries the penalty 100 E2 - 100 E_p E_m instead (QED_model.py:131-141, 381-393)
        self.link_symmetries = bool(link_symmetries)
        self.get_only_bulk = bool(get_only_bulk)
        charges = [0] * self.n_sites
        if self.bg_list is not None:
            charges = [int(round(q)) for q in self.bg_list]
            for i, q in enumerate(self.bg_list):
                if abs(q - charges[i]) > 1e-12:
                    raise ValueError(f"No GI states at site {i} with bg {q}")
        site = opx.qed_site(self.spin, self.pure_theory, self.dim, backgrounds=set(charges))
        labels = []
        for i in range(self.n_sites):
            base = self.lattice.label(i, self.staggered_basis)
            lab = opx.bg_label(base, charges[i])
            if lab not in site.states:
                raise ValueError(f"No GI states at site {i} label={base} with bg {float(charges[i])}")
            labels.append(lab)
        with_bg = self.background > 0
        self._project(site, lambda l: opx.qed_operators(site, self.spin, self.pure_theory, self.dim, l,
                                                        with_bg=with_bg), labels)
        if with_bg:
            # the reference's per-site gauge states lead with the background value
            self.gauge_states_per_site = [np.column_stack([np.full(v.shape[0], q), v])
                                          for q, v in zip(charges, self.gauge_states_per_site)]
        d_max = self.ops["E_px"].shape[1]

        def diag(name):
            return np.real(np.einsum("sii->si", self.ops[name]))

        glob = None
        if not self.pure_theory:
            glob = (diag("N")[None], np.array([float(self.n_sites // 2)]), "U")
        pairs = self.lattice.link_pairs()
        link = (np.stack([np.stack([diag(f"E_p{d}"), -diag(f"E_m{d}")]) for d in self.directions]),
                np.zeros(self.dim), pairs)
        nbody = None
        if self.pure_theory and not any(self.has_obc) and self.dim == 2:
            line_of = {"x": "y", "y": "x"}
            nbody = (np.stack([diag(f"E_p{d}") for d in self.directions]), np.zeros(self.dim),
                     [self.lattice.origin_line(line_of[d]) for d in self.directions], "U")
        elif self.pure_theory and not any(self.has_obc) and self.dim == 3:
            # electric flux through the three coordinate planes at the origin
            nbody = (np.stack([diag(f"E_p{d}") for d in self.directions]), np.zeros(3),
                     [self.lattice.origin_face(d) for d in self.directions], "U")
        self.basis = SectorBasis(self.loc_dims, glob=glob, link=link if self.link_symmetries else None,
                                 nbody=nbody)
        logger.info(f"SEC DIM: {self.basis.n_states}")
        self.default_params()

    def _init_integrated(self):
        """1D QED with the gauge fields integrated out: two states per site
        (Pauli basis, index 0 = occupied, QED_operators.py:541-545 with
        spin_operators.py:62-72), sector N_tot = n_sites // 2, no link
        constraints (QED_model.py:64-82, 114-140)."""
        self.spin = "integrated"
        self.pure_theory = False
        self.plaq_basis = False
        self.link_symmetries = False
        self.staggered_basis = False
        ops = _integrated_operators()
        site = _PauliSite()
        self._project(site, lambda l: ops, ["site"] * self.n_sites)
        glob = (np.real(np.einsum("sii->si", self.ops["N"]))[None], np.array([float(self.n_sites // 2)]), "U")
        self.basis = SectorBasis(self.loc_dims, glob=glob)
        logger.info(f"SEC DIM: {self.basis.n_states}")
        self.default_params()

    def _integrated_static_charges(self):
        """QED_model.py:494-503."""
        if self.bg_list is None:
            return np.zeros(self.n_sites)
        return np.asarray(self.bg_list, dtype=float)

    def _build_integrated(self, g, m, theta, dtype_mode):
        """build_integrated_Hamiltonian (QED_model.py:396-492): staggered mass,
        hopping S+ S- with t = -i/2, and the electric energy
        E sum_{n<L-1} [sum_{k<=n} (q_k + (Sz_k + (-1)^k)/2)]^2 expanded into
        one-body Sz terms h_i, two-body Sz_i Sz_j terms and a constant."""
        mode = dtype_mode
        if mode in ("auto", None):
            mode = "complex"
        real = self._dtype_real(theta, mode)
        if m is None:
            raise ValueError("m is required with matter fields")
        self.QED_Hamiltonian_couplings(g, m, theta)
        self._new_program()
        self._integrated_matter(self.coeffs)
        self._integrated_electric(self.coeffs["E"])
        return self._build_program(real)

    def _integrated_matter(self, c):
        """Staggered mass and hopping placements (QED_model.py:409-427)."""
        lat = self.lattice
        even = lat.even_mask()
        self._add_local("N", c["m_even"], [i for i in range(self.n_sites) if even[i]])
        self._add_local("N", c["m_odd"], [i for i in range(self.n_sites) if not even[i]])
        self._add_nbody(["Sp", "Sm"], lat.bonds("x"), c["t"], True)

    def _integrated_electric(self, E):
        """Electric energy placements at coupling E (QED_model.py:428-491):
        h_i Sz_i with h_i = E sum_{i<=n<L-1} A_n (A the prefix sums of
        q_k + (-1)^k / 2), J_ij Sz_i Sz_j with J_ij = E (L-1-j) / 2 for
        i < j < L-1, and the constant E [sum_n A_n^2 + n_links (n_links+1) / 8]
        as a diagonal term on site 0."""
        n = self.n_sites
        n_links = n - 1
        q = self._integrated_static_charges()
        offsets = q + 0.5 * np.array([1.0 if i % 2 == 0 else -1.0 for i in range(n)])
        prefix = np.cumsum(offsets)
        for k in range(n_links):
            h = E * np.sum(prefix[k:n_links])
            if abs(h) >= 1e-14:
                self._add_local("Sz", h, [k])
        for j in range(1, n_links):
            w = 0.5 * E * (n - 1 - j)
            if abs(w) > 1e-14:
                self._add_nbody(["Sz", "Sz"], [(i, j) for i in range(j)], w, False)
        if n_links > 0:
            const = E * (np.sum(prefix[:n_links] ** 2) + 0.25 * 0.5 * n_links * (n_links + 1))
            if abs(const) > 1e-14:
                self._add_local("I", float(const), [0])

    def reconstruct_integrated_E2_from_N(self, density_obs_name="N", density_corr_obs_name="N_N",
                                         state_index=None, dynamics=False, compute_density_corr=True,
                                         print_values=True):
        """Link-resolved <E^2> of integrated 1D QED from the measured density
        (QED_model.py:505-652): E_n = sum_{k<=n} [q_k + N_k + ((-1)^k - 1)/2],
        so <E_n^2> = B_n^2 + 2 B_n sum_{k<=n} <N_k> + sum_{k,l<=n} <N_k N_l>
        with B_n = sum_{k<=n} [q_k + ((-1)^k - 1)/2].  The correlator is read
        from res or measured on the device."""
        if self.spin != "integrated" or self.dim != 1 or self.pure_theory:
            raise ValueError("Integrated E2 reconstruction is defined only for 1D QED with matter")
        if density_obs_name not in self.res:
            raise KeyError(f"Missing '{density_obs_name}' in self.res. Measure local density first.")
        dens = np.asarray(self.res[density_obs_name], dtype=float)
        if dens.shape != (self.n_sites,):
            raise ValueError(f"Expected {density_obs_name}.shape=({self.n_sites},) got {dens.shape}.")
        n_links = self.n_sites - 1
        if n_links <= 0:
            self.res["E2"] = np.zeros(0)
            self.res["E2_avg"] = 0.0
            return self.res["E2"]
        if density_corr_obs_name in self.res:
            corr = np.asarray(self.res[density_corr_obs_name], dtype=float)
        else:
            if not compute_density_corr:
                raise KeyError(f"Missing '{density_corr_obs_name}' in self.res. "
                               "Measure ['N','N_N'] or set compute_density_corr=True.")
            states = self.H.psi_time if dynamics else getattr(self.H, "Npsi", None)
            if states is None:
                raise ValueError("Cannot compute density correlator: no states available.")
            if state_index is None:
                if len(states) != 1:
                    raise ValueError("state_index is required to compute density correlator "
                                     "when several states are available.")
                state_index = 0
            psi_dev = self._sector_state(states[state_index].psi_device)
            saved = getattr(self, "_obs_groups", {})
            pairs = [(i, j) for i in range(self.n_sites) for j in range(i + 1, self.n_sites)]
            self._obs_groups = {"N_N": ("twobody", ["N", "N"], pairs)}
            try:
                v = np.real(self._group_expvals(psi_dev)["N_N"])
            finally:
                self._obs_groups = saved
            corr = np.zeros((self.n_sites, self.n_sites))
            for (i, j), x in zip(pairs, v):
                corr[i, j] = corr[j, i] = x
            self.res[density_corr_obs_name] = corr
        if corr.shape != (self.n_sites, self.n_sites):
            raise ValueError(f"Expected '{density_corr_obs_name}' shape ({self.n_sites}, {self.n_sites}), "
                             f"got {corr.shape}.")
        corr = np.array(corr, dtype=float, copy=True)
        np.fill_diagonal(corr, dens)  # N_k^2 = N_k for fermions
        q = self._integrated_static_charges()
        k = np.arange(self.n_sites, dtype=float)
        b = np.cumsum(q + 0.5 * ((-1.0) ** k - 1.0))
        pd = np.cumsum(dens)
        pc = np.diag(corr.cumsum(axis=0).cumsum(axis=1))
        e2 = b[:n_links] ** 2 + 2.0 * b[:n_links] * pd[:n_links] + pc[:n_links]
        self.res["E2"] = e2
        self.res["E2_avg"] = float(np.mean(e2))
        if print_values:
            logger.info("----------------------------------------------------")
            logger.info("E2 ")
            for i in range(n_links):
                logger.info(f"{(i,)} {format(e2[i], '.16f')}")
            logger.info(f"{format(self.res['E2_avg'], '.16f')}")
        return e2

    def check_symmetries(self):
        """E_p{d}(i) - E_m{d}(i+d) = 0 on every link, from the measured fields
        (QED_model.py:707-717; measured by get_observables([... "E_p{d}",
        "E_m{d}" ...]) and measure_observables).  The relation checked is the
        one the sector imposes ([E_p, -E_m] summing to 0, QED_model.py:136);
        the reference's check, which reads LocalTerm objects it no longer
        keeps, passes the opposite sign."""
        for d in self.directions:
            self.check_link_symmetry(d, f"E_p{d}", f"E_m{d}", value=0, sign=-1)

    def QED_Hamiltonian_couplings(self, g, m=None, theta=0.0):
        """QED_model.py:719-776 (dressed-site normalisation)."""
        c = {"g": g, "E": g / 2, "B": -1 / (2 * g), "theta": -theta * g}
        if not self.pure_theory and m is not None:
            t = 0.5
            c.update({"t": -complex(0, t), "tx_even": -complex(0, t), "tx_odd": -complex(0, t),
                      "ty_even": -t, "ty_odd": t, "tz_even": -complex(0, t), "tz_odd": complex(0, t),
                      "m": m, "m_odd": -m, "m_even": m})
        self.coeffs = c

    def _dtype_real(self, theta, dtype_mode):
        mode = dtype_mode
        if isinstance(mode, (bool, np.bool_)):
            mode = "complex" if mode else "real"
        if mode is None or mode == "auto":
            return self.pure_theory and abs(theta) <= 1e-12
        if mode not in ("real", "complex"):
            raise ValueError(f"dtype_mode must be one of ['auto', 'complex', 'real'], got {dtype_mode!r}")
        return mode == "real"

    def build_Hamiltonian(self, g, m=None, theta=0.0, dtype_mode="auto"):
        """build_truncated_Hamiltonian term order (QED_model.py:235-394), or
        build_integrated_Hamiltonian for spin='integrated'."""
        if self.spin == "integrated":
            return self._build_integrated(g, m, theta, dtype_mode)
        real = self._dtype_real(theta, dtype_mode)
        self.QED_Hamiltonian_couplings(g, m, theta)
        c = self.coeffs
        lat = self.lattice
        self._new_program()
        self._add_local("E2", c["E"])
        if self.dim > 1:
            names = ["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"]
            plaq = lat.plaquettes(("x", "y"))
            self._add_nbody(names, plaq, c["B"], True)
        if self.dim == 3:
            self._add_nbody(["C_px,pz", "C_pz,mx", "C_mz,px", "C_mx,mz"], lat.plaquettes(("x", "z")), c["B"], True)
            self._add_nbody(["C_py,pz", "C_pz,my", "C_mz,py", "C_my,mz"], lat.plaquettes(("y", "z")), c["B"], True)
        if self.dim == 2 and abs(c["theta"]) > 10e-10:
            self._add_nbody(["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"], lat.plaquettes(("x", "y")),
                            c["theta"], True)
        if self.dim == 3 and abs(c["theta"]) > 10e-10:
            self._add_nbody(["EzC_px,py", "C_py,mx", "C_my,px", "C_mx,my"], lat.plaquettes(("x", "y")), c["theta"], True)
            self._add_nbody(["EyC_px,pz", "C_pz,mx", "C_mz,px", "C_mx,mz"], lat.plaquettes(("x", "z")), c["theta"], True)
            self._add_nbody(["ExC_py,pz", "C_pz,my", "C_mz,py", "C_my,mz"], lat.plaquettes(("y", "z")), c["theta"], True)
        if not self.pure_theory:
            if m is None:
                raise ValueError("m is required with matter fields")
            even = lat.even_mask()
            self._add_local("N", c["m_even"], [i for i in range(self.n_sites) if even[i]])
            self._add_local("N", c["m_odd"], [i for i in range(self.n_sites) if not even[i]])
            for d in self.directions:
                for stag in ("even", "odd"):
                    bonds = [b for b in lat.bonds(d) if even[b[0]] == (stag == "even")]
                    self._add_nbody([f"Q_p{d}_dag", f"Q_m{d}"], bonds, c[f"t{d}_{stag}"], True)
        if not self.link_symmetries:
            # link penalties: 100 sum E2 - 100 sum_links E_p(i) E_m(i + d)
            self._add_local("E2", 100.0)
            for d in self.directions:
                self._add_nbody([f"E_p{d}", f"E_m{d}"], lat.bonds(d), -100.0, False)
        return self._build_program(real)

    def _gauge_weights(self):
        """Staggered fermions hop with the coefficients -i/2 along x, -i/2 and
        +i/2 along z and real ones along y (QED_Hamiltonian_couplings, from
        QED_model.py:719-776), and every hopping entry moves one fermion
        between the bond's sites.  With a_j = (x_j + z_j) mod 2 the phase
        i^(sum_j a_j N_j) therefore turns each hopping entry real; the electric,
        mass and plaquette terms keep every N_j and stay real.  A lattice
        periodic along x or z with an odd extent has a wrap-around bond whose
        ends share a_j: its entries stay imaginary and real_gauge_matrix
        refuses the real build."""
        if self.spin == "integrated" or self.pure_theory:
            return None
        a = np.array([(c[0] + (c[2] if self.dim == 3 else 0)) % 2
                      for c in map(self.lattice.coords, range(self.n_sites))], dtype=np.int64)
        return a, "N"

    def build_truncated_Hamiltonian(self, g, m=None, theta=0.0, dtype_mode="auto"):
        """The truncated-link builder by name (QED_model.py:235-394);
        build_Hamiltonian dispatches to it for integer spins."""
        if self.spin == "integrated":
            raise ValueError("this model integrates the gauge fields out: use build_integrated_Hamiltonian")
        return self.build_Hamiltonian(g, m, theta, dtype_mode)

    def build_integrated_Hamiltonian(self, g, m=None, theta=0.0, dtype_mode="auto"):
        """The integrated-gauge builder by name (QED_model.py:396-652), for
        spin='integrated' models."""
        if self.spin != "integrated":
            raise ValueError("this model has truncated links: use build_truncated_Hamiltonian")
        return self.build_Hamiltonian(g, m, theta, dtype_mode)

    def overlap_QMB_state(self, name):
        """Named reference configurations (QED_model.py:778-899), one local-state
        index per site, for get_qmb_state_from_configs:
          1D with matter, spin 1: "V" (bare vacuum: every link E = 0, even sites
            empty, odd sites filled; alias "PV") and "meson_center" (the central
            (even, odd) pair turned into a meson: the even site filled with its
            right link at E = +1); the same two for spin='integrated' as Pauli
            occupations;
          3D, 2x2x2 with static charges bg_list = [-1, 0, 0, 0, 0, 0, 0, 1],
            spin 1, 2 or 3: the minimal flux strings "S1".."S6"."""
        if self.dim == 1 and not self.pure_theory:
            key = name.strip().lower().replace(" ", "_")
            if key == "pv":
                key = "v"
            if key in ("v", "meson_center") and self.spin == "integrated":
                # occupations in the Pauli basis (index 0 = occupied): even sites
                # empty, odd sites filled; the meson flips the central (even, odd)
                # pair (QED_model.py:803-818)
                n = self.n_sites
                cfg = np.array([1 if i % 2 == 0 else 0 for i in range(n)], dtype=np.int32)
                if key == "meson_center":
                    even = n // 2
                    if even % 2 == 1:
                        even -= 1
                    even = max(0, min(even, n - 2))
                    cfg[even], cfg[even + 1] = 0, 1
                return cfg
            if key in ("v", "meson_center"):
                if self.spin != 1:
                    raise NotImplementedError("1D labels 'V'/'meson_center' are implemented for "
                                              "spin='integrated' or truncated spin=1.")
                n = self.n_sites
                # spin-1 1D states in product order over (n, E_mx, E_px): index 1
                # of an even site is (0, 0, 0), index 3 of an odd site (1, 0, 0);
                # on an OBC edge the kept states are those with a zero outer link
                cfg = np.array([1 if i % 2 == 0 else 3 for i in range(n)], dtype=np.int32)
                if all(self.has_obc):
                    cfg[0], cfg[n - 1] = 0, 1
                if key == "meson_center":
                    even = n // 2
                    if even % 2 == 1:
                        even -= 1
                    even = max(0, min(even, n - 2))
                    cfg[even], cfg[even + 1] = 4, 1
                return cfg
        if self.dim == 3 and self.bg_list == [-1, 0, 0, 0, 0, 0, 0, 1]:
            if self.spin not in _QED3D_STRINGS:
                raise NotImplementedError("3D string states are tabulated for spin 1, 2 and 3")
            if name not in _QED3D_STRINGS[self.spin]:
                raise ValueError(f"Unknown String name: {name}")
            return np.array(_QED3D_STRINGS[self.spin][name])
        raise NotImplementedError("Only 3D QED states are supported")

    def multi_term_program(self, m=None):
        """(tables, placements with term ids) of H(g) = (g/2) H_E + (-1/(2g)) H_B + H_0
        for a g sweep at fixed m: term 0 = E2, term 1 = xy plaquettes (+ h.c.; the
        3D planes too), term 2 = mass + hopping at mass m (empty without matter)
        and the link penalties when link_symmetries=False.
        Coefficients of a coupling g: QED_Model.multi_term_coefficients(g).
        spin='integrated': term 0 = the electric energy at E = 1, term 1
        empty, term 2 = mass + hopping."""
        lat = self.lattice
        self._new_program()
        if self.spin == "integrated":
            if m is None:
                raise ValueError("m is required with matter fields")
            self._integrated_electric(1.0)
            for p in self._placements:
                p["term"] = 0
            k0 = len(self._placements)
            self.QED_Hamiltonian_couplings(1.0, m)
            self._integrated_matter(self.coeffs)
            for p in self._placements[k0:]:
                p["term"] = 2
            return np.stack(self._tables), list(self._placements)
        self._add_local("E2", 1.0)
        for p in self._placements:
            p["term"] = 0
        k0 = len(self._placements)
        if self.dim > 1:
            self._add_nbody(["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"], lat.plaquettes(("x", "y")), 1.0, True)
        if self.dim == 3:
            self._add_nbody(["C_px,pz", "C_pz,mx", "C_mz,px", "C_mx,mz"], lat.plaquettes(("x", "z")), 1.0, True)
            self._add_nbody(["C_py,pz", "C_pz,my", "C_mz,py", "C_my,mz"], lat.plaquettes(("y", "z")), 1.0, True)
        for p in self._placements[k0:]:
            p["term"] = 1
        k1 = len(self._placements)
        if not self.pure_theory:
            if m is None:
                raise ValueError("m is required with matter fields")
            self.QED_Hamiltonian_couplings(1.0, m)
            c = self.coeffs
            even = lat.even_mask()
            self._add_local("N", c["m_even"], [i for i in range(self.n_sites) if even[i]])
            self._add_local("N", c["m_odd"], [i for i in range(self.n_sites) if not even[i]])
            for d in self.directions:
                for stag in ("even", "odd"):
                    bonds = [b for b in lat.bonds(d) if even[b[0]] == (stag == "even")]
                    self._add_nbody([f"Q_p{d}_dag", f"Q_m{d}"], bonds, c[f"t{d}_{stag}"], True)
        if not self.link_symmetries:
            self._add_local("E2", 100.0)
            for d in self.directions:
                self._add_nbody([f"E_p{d}", f"E_m{d}"], lat.bonds(d), -100.0, False)
        for p in self._placements[k1:]:
            p["term"] = 2
        return np.stack(self._tables), list(self._placements)


    def build_multi_term(self, m=None, dtype_mode="auto", real_gauge=False):
        """Assembles the three coupling-independent parts of H(g) once on one
        pattern; the matrix at coupling g is sum_t multi_term_coefficients(g)[t] H_t
        (products: CSRMatrix.matvec_coef, Server).  The couplings of the last
        build_Hamiltonian (self.coeffs) are kept.  real_gauge=True assembles
        the parts of D^H H(g) D instead (the term coefficients are real, so one
        phase per placement serves every g; see real_gauge_matrix) with float64
        values: the same eigenvalues, eigenvectors in the gauge of D
        (csr.real_gauge says whether the real build applied)."""
        saved = getattr(self, "coeffs", None)
        try:
            tables, placements = self.multi_term_program(m)
        finally:
            if saved is not None:
                self.coeffs = saved
        if real_gauge:
            rg = self._real_gauge_placements((list(tables), placements))
            if rg is not None:
                try:
                    csr = CSRMatrix(self.basis, tables, rg, L.VALUE_F64, n_terms=3)
                    csr.real_gauge = True
                    return csr
                except ValueError:
                    logger.info("no real gauge: D^H H D has complex entries")
        real = self._dtype_real(0.0, dtype_mode)
        csr = CSRMatrix(self.basis, tables, placements, L.VALUE_F64 if real else L.VALUE_C128, n_terms=3)
        csr.real_gauge = False
        return csr


    def multi_term_coefficients(self, g):
        """Term coefficients of H(g) (QED_Hamiltonian_couplings, QED_model.py:719-776)."""
        return [g / 2, -1 / (2 * g), 1.0]


# Minimal string configurations of 3D QED (2x2x2 OBC, matter, static charges
# -1 at site 0 and +1 at site 7) per link truncation: local-state indices,
# one per site (QED_model.py:855-897).  Data that names states of the
# reference's basis order, kept as the reference defines it.
_QED3D_STRINGS = {
    1: {"S1": [4, 9, 9, 3, 6, 0, 3, 10], "S2": [3, 9, 6, 0, 9, 3, 3, 11], "S3": [1, 7, 9, 3, 9, 2, 3, 10],
        "S4": [4, 9, 9, 3, 7, 3, 0, 8], "S5": [1, 6, 9, 2, 9, 3, 3, 11], "S6": [3, 9, 7, 3, 9, 3, 2, 8]},
    2: {"S1": [11, 27, 27, 9, 22, 4, 9, 29], "S2": [10, 27, 22, 4, 27, 9, 9, 30],
        "S3": [10, 27, 23, 9, 27, 9, 8, 25], "S4": [6, 22, 27, 8, 27, 9, 9, 30],
        "S5": [6, 23, 27, 9, 27, 8, 9, 29], "S6": [11, 27, 27, 9, 23, 9, 4, 25]},
    3: {"S1": [20, 54, 47, 11, 54, 18, 18, 58], "S2": [21, 54, 54, 18, 47, 11, 18, 57],
        "S3": [21, 54, 54, 18, 48, 18, 11, 51], "S4": [20, 54, 48, 18, 54, 18, 17, 51],
        "S5": [14, 48, 54, 18, 54, 17, 18, 57], "S6": [14, 47, 54, 17, 54, 18, 18, 58]},
}


class Zn_Model(QuantumModel):
    """Pure Z_n lattice gauge theory on a 2D lattice (no model class exists in
    the reference; built from its Z_n operator factories,
    edlgt/operators/Zn_operators.py:130-360, in the shape of QED_Model).

    Sector: Gauss law in the local basis, link constraints
    E_p{d}(i) + E_m{d}(i+d) = 0, and for PBC optionally the electric flux through
    the origin lines fixed modulo n (flux_sector=(fx, fy); None keeps all flux
    sectors).  H = -J sum_plaq (U_plaq + h.c.) - h sum_links (E_l + E_l^dag)."""

    def __init__(self, n, pure_theory=True, flux_sector=(0, 0), **kwargs):
        super().__init__(**kwargs)
        if not pure_theory:
            raise NotImplementedError("Z_n with matter is not on the device path")
        if self.dim != 2:
            raise NotImplementedError("Z_n models are 2D")
        self.n = int(n)
        self.pure_theory = True
        site = opx.zn_site(self.n)
        labels = [self.lattice.label(i, False) for i in range(self.n_sites)]
        self._project(site, lambda l: opx.zn_operators(site, self.n, l), labels)

        def diag(name):
            return np.real(np.einsum("sii->si", self.ops[name]))

        link = (np.stack([np.stack([diag(f"E_p{d}"), diag(f"E_m{d}")]) for d in self.directions]),
                np.zeros(self.dim), self.lattice.link_pairs())
        nbody = None
        self.flux_sector = None
        if flux_sector is not None and not any(self.has_obc):
            if self.n % 2 == 0:
                raise NotImplementedError("flux sectors are defined for odd n (integer E)")
            line_of = {"x": "y", "y": "x"}
            nbody = (np.stack([diag(f"E_p{d}") for d in self.directions]),
                     np.array(flux_sector, dtype=float),
                     [self.lattice.origin_line(line_of[d]) for d in self.directions], L.SYM_MOD(self.n))
            self.flux_sector = tuple(flux_sector)
        self.basis = SectorBasis(self.loc_dims, link=link, nbody=nbody)
        self.default_params()

    def build_Hamiltonian(self, J, h, dtype_mode="auto"):
        self.coeffs = {"J": J, "h": h}
        self._new_program()
        self._add_local("E_clock", -h)
        self._add_nbody(["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"], self.lattice.plaquettes(("x", "y")),
                        -J, True)
        real = dtype_mode in ("auto", "real", False, None)
        return self._build_program(real)

    def multi_term_program(self):
        """(tables, placements with term ids) of H = J*H_J + h*H_h for batched solves:
        term 0 = -sum_plaq(U + h.c.), term 1 = -sum_links(E + E^dag)."""
        self._new_program()
        self._add_local("E_clock", -1.0)
        for p in self._placements:
            p["term"] = 1
        k0 = len(self._placements)
        self._add_nbody(["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"], self.lattice.plaquettes(("x", "y")),
                        -1.0, True)
        for p in self._placements[k0:]:
            p["term"] = 0
        return np.stack(self._tables), list(self._placements)

    def build_multi_term(self):
        """Assembles H_J and H_h once on one shared pattern (real CSR with
        term ids): H(J, h) = J * H_J + h * H_h for any couplings, without
        re-assembly.  Returns the CSRMatrix (coefficients [J, h])."""
        tables, placements = self.multi_term_program()
        return CSRMatrix(self.basis, tables, placements, L.VALUE_F64, n_terms=2)


class SU2_Model(QuantumModel):
    """SU(2) lattice gauge theory with optional two-colour staggered matter
    (edlgt/models/SU2_model.py:28-355, 357-484, 486-624), 1D to 3D: the
    hard-coded spin-1/2 operator set, or the generalised one (any rishon
    truncation spin, use_generic_model=True or spin >= 1).

    Sector: global N_tot = sectors[0] with matter, link constraints
    T2_p{d}(i) = T2_m{d}(i+d), and for the pure theory on periodic lattices
    the parity of the flux through the origin lines (2D) or faces (3D) fixed
    to +1."""
    _pad_state_values = False  # SU(2) prints the local values as they are

    def __init__(self, spin, pure_theory, bg_list=None, sectors=None, use_generic_model=False, **kwargs):
        super().__init__(**kwargs)
        s2 = 2 * float(spin)
        if s2 < 1 or abs(s2 - round(s2)) > 1e-12:
            raise ValueError(f"spin must be a positive multiple of 1/2, not {spin}")
        # static background charges (SU2_model.py:63-64, 100-104): one spin
        # per site, the sites' bases restricted to their charge
        background = Fraction(0)
        self.bg_list = None
        if bg_list is not None:
            bgv = np.asarray(bg_list, dtype=float)
            if bgv.shape != (self.n_sites,):
                raise ValueError(f"len(bg_list) must match n_sites = {self.n_sites}: got {bgv.shape}")
            if np.any(bgv < 0) or np.any(np.abs(2 * bgv - np.round(2 * bgv)) > 1e-12):
                raise ValueError("background charges are spins: non-negative multiples of 1/2")
            background = Fraction(int(round(2 * bgv.max())), 2)
            if background > 0:
                self.bg_list = bgv.tolist()
        if self.dim > 3:
            raise ValueError("SU(2) models are defined in 1D, 2D and 3D")
        from .su2_operators import SU2Site
        s_max = Fraction(int(round(s2)), 2)
        generic = bool(use_generic_model) or s_max >= 1
        if self.dim == 3 and s_max > Fraction(1, 2):
            raise NotImplementedError("3D SU(2) is built for spin 1/2 rishons (a spin-1 site has "
                                      "2,985,984 states before projection)")
        self.spin = float(spin)
        self.pure_theory = bool(pure_theory)
        self.use_generic_model = generic
        # translation-combined symmetry of the 1D chain with matter: local
        # involution on the six dressed states and the sign of each state
        # (translational_sym.py:1013-1014, 1171-1172)
        if not pure_theory and self.dim == 1 and not generic:
            self.tc_symmetry_tables = (np.array([4, 5, 2, 3, 0, 1]), np.array([1.0, 1.0, 1.0, -1.0, 1.0, 1.0]))
        self.background = float(background)
        site = SU2Site(self.pure_theory, self.dim, s_max, generic, background)
        if max(site.dim(l) for l in site.basis) > 255:
            raise NotImplementedError(f"a dressed site of {max(site.dim(l) for l in site.basis)} states "
                                      "exceeds the engine's 255 local states")
        self._site = site
        labels = []
        for i in range(self.n_sites):
            lab = self.lattice.label(i, False)
            if self.bg_list is not None:
                lab = f"{lab}@bg{int(round(2 * self.bg_list[i]))}"
            if lab not in site.basis:
                raise ValueError(f"No GI states at site {i} label={lab}")
            labels.append(lab)
        self._project(site, site.operators, labels)

        def diag(name):
            return np.real(np.einsum("sii->si", self.ops[name]))

        glob = None
        if not self.pure_theory:
            if sectors is None:
                raise ValueError("sectors=[N_tot] is required with matter fields")
            glob = (diag("N_tot")[None], np.array([float(sectors[0])]), "U")
            self.zero_density = sectors[0] == self.n_sites
        link = (np.stack([np.stack([diag(f"T2_p{d}"), -diag(f"T2_m{d}")]) for d in self.directions]),
                np.zeros(self.dim), self.lattice.link_pairs())
        nbody = None
        if self.pure_theory and not any(self.has_obc) and self.dim == 2:
            line_of = {"x": "y", "y": "x"}
            nbody = (np.stack([diag(f"P_p{d}") for d in self.directions]), np.ones(self.dim),
                     [self.lattice.origin_line(line_of[d]) for d in self.directions], "Z")
        elif self.pure_theory and not any(self.has_obc) and self.dim == 3:
            nbody = (np.stack([diag(f"P_p{d}") for d in self.directions]), np.ones(3),
                     [self.lattice.origin_face(d) for d in self.directions], "Z")
        self.basis = SectorBasis(self.loc_dims, glob=glob, link=link, nbody=nbody)
        if self.basis.n_states == 0:
            raise ValueError("No configurations found for the given symmetry sectors")
        logger.info(f"SEC DIM: {self.basis.n_states}")
        self.default_params()

    def _gauge_weights(self):
        """The hopping coefficients are -2 sqrt(2) i along x and real along y
        and z (SU2_Hamiltonian_couplings, from SU2_model.py:555-624), and every
        hopping entry moves one fermion across its bond: with a_j = x_j mod 2
        the phase i^(sum_j a_j N_tot_j) turns the x entries real; the other
        terms keep every N_tot_j.  A chain or lattice periodic along x with an
        odd extent keeps imaginary wrap-around entries (real build refused)."""
        if self.pure_theory:
            return None
        a = np.array([c[0] % 2 for c in map(self.lattice.coords, range(self.n_sites))], dtype=np.int64)
        return a, "N_tot"

    def check_symmetries(self):
        """T2_p{d}(i) - T2_m{d}(i+d) = 0 on every link, from the measured
        Casimirs (SU2_model.py:474-484)."""
        for d in self.directions:
            self.check_link_symmetry(d, f"T2_p{d}", f"T2_m{d}", value=0, sign=-1)

    def SU2_Hamiltonian_couplings(self, g, m=None, theta=0.0):
        """SU2_model.py:555-624 (hopping rescaled to 2 sqrt(2))."""
        if self.dim == 1:
            E, B = 8 * g / 3, 0
        else:
            E, B = g / 2, -1 / (2 * g)
        c = {"g": g, "E": E, "B": B, "theta": -theta * g}
        if not self.pure_theory and m is not None:
            t = 2 * np.sqrt(2)
            c.update({"tx_even": -complex(0, t), "tx_odd": -complex(0, t), "ty_even": -t, "ty_odd": t,
                      "tz_even": -t, "tz_odd": t, "m": m, "m_odd": -m, "m_even": m})
        self.coeffs = c

    def build_Hamiltonian(self, g, m=None, theta=0.0, lambda_noise=0.0, dtype_mode="auto"):
        """build_base_Hamiltonian term order (SU2_model.py:200-355)."""
        mode = dtype_mode
        if isinstance(mode, (bool, np.bool_)):
            mode = "complex" if mode else "real"
        if mode in (None, "auto"):
            real = self.pure_theory and abs(theta) <= 1e-12
        elif mode in ("real", "complex"):
            real = mode == "real"
        else:
            raise ValueError(f"dtype_mode must be one of ['auto', 'complex', 'real'], got {dtype_mode!r}")
        # the Gauss-law-violating noise term needs background charges and periodic
        # boundaries (SU2_model.py:338-354); without background charges the
        # reference adds nothing, as here
        self.SU2_Hamiltonian_couplings(g, m, theta)
        if abs(self.coeffs["theta"]) > 10e-10 and self.dim == 3 and self.use_generic_model:
            raise NotImplementedError("the theta term is built with the hard-coded spin-1/2 operators")
        if self.use_generic_model:
            return self._build_generic(m, real)
        c = self.coeffs
        lat = self.lattice
        self._new_program()
        self._add_local("E_square", c["E"])
        if self.dim > 1:
            self._add_nbody(["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"], lat.plaquettes(("x", "y")), c["B"], True)
        if self.dim == 3:
            self._add_nbody(["C_px,pz", "C_pz,mx", "C_mz,px", "C_mx,mz"], lat.plaquettes(("x", "z")), c["B"], True)
            self._add_nbody(["C_py,pz", "C_pz,my", "C_mz,py", "C_my,mz"], lat.plaquettes(("y", "z")), c["B"], True)
        if not self.pure_theory:
            if m is None:
                raise ValueError("m is required with matter fields")
            even = lat.even_mask()
            self._add_local("N-1", c["m_even"], [i for i in range(self.n_sites) if even[i]])
            self._add_local("N-1", c["m_odd"], [i for i in range(self.n_sites) if not even[i]])
            for d in self.directions:
                for stag in ("even", "odd"):
                    bonds = [b for b in lat.bonds(d) if even[b[0]] == (stag == "even")]
                    self._add_nbody([f"Qp{d}_dag", f"Qm{d}"], bonds, c[f"t{d}_{stag}"], True)
        if self.dim == 3 and abs(c["theta"]) > 10e-10:
            # topological term: E_normal times the plaquette, per plane (SU2_model.py:303-336)
            for names, axes in ((["EzC_px,py", "C_py,mx", "C_my,px", "C_mx,my"], ("x", "y")),
                                (["EyC_px,pz", "C_pz,mx", "C_mz,px", "C_mx,mz"], ("x", "z")),
                                (["ExC_py,pz", "C_pz,my", "C_mz,py", "C_my,mz"], ("y", "z"))):
                self._add_nbody(names, lat.plaquettes(axes), c["theta"], True)
        if self.background > 0 and lambda_noise > 0 and not any(self.has_obc):
            # Gauss-law-violating hopping through the background charges (SU2_model.py:338-354)
            if "Vpx_dag" not in self.ops:
                raise NotImplementedError("the noise term exists for the 1D chain with matter and "
                                          "spin-1/2 background charges")
            self._add_nbody(["Vpx_dag", "Vmx"], lat.bonds("x"), -lambda_noise, True)
        return self._build_program(real)

    def build_local_Hamiltonian(self, g, m, R0):
        """The energy density around site R0 of the 1D chain with matter
        (SU2_model.py:626-699), stored in self.Hlocal: the electric term and
        the staggered mass (-1)^R0 m at R0, and half of each hopping on the
        bonds (R0, R0+1) and (R0-1, R0) with its conjugate written out as the
        reference writes it ([Qpx_dag, Qmx] times t/2, [Qpx, Qmx_dag] times
        -t/2).  Summed over R0 it is the chain's Hamiltonian.  After
        set_momentum_pair, Hlocal holds the rectangular block P_kL^dagger H_R0
        P_kR (SciPy CSR) instead."""
        from .device import CSRMatrix
        if self.dim > 1:
            raise ValueError(f"Local Hamiltonian valid only for D=1, got {self.dim}")
        if self.pure_theory or self.use_generic_model:
            raise ValueError("the local Hamiltonian is defined for the spin-1/2 chain with matter")
        if self.momentum_basis is not None:
            raise NotImplementedError("the local Hamiltonian is built in the sector basis or, after "
                                      "set_momentum_pair, between two momentum sectors")
        if not 0 <= int(R0) < self.n_sites:
            raise ValueError(f"R0 must be a site of the chain, got {R0}")
        R0 = int(R0)
        # the working program and couplings of the last build stay as they were
        saved = (self._tables, self._table_ids, self._placements, dict(self.coeffs)) \
            if hasattr(self, "_placements") and hasattr(self, "coeffs") else None
        try:
            self.SU2_Hamiltonian_couplings(g, m)
            c = self.coeffs
            self._new_program()
            self._add_local("E_square", c["E"], [R0])
            self._add_local("N-1", ((-1) ** R0) * c["m"], [R0])
            half = 0.5 * c["tx_even"]
            fwd = [b for b in self.lattice.bonds("x") if b[0] == R0]
            bwd = [b for b in self.lattice.bonds("x") if b[0] == (R0 - 1) % self.n_sites]
            for names, strength in ((["Qpx_dag", "Qmx"], half), (["Qpx", "Qmx_dag"], -half)):
                self._add_nbody(names, fwd, strength, False)
                self._add_nbody(names, bwd, strength, False)
            csr = CSRMatrix(self.basis, np.stack(self._tables), self._placements, L.VALUE_C128)
        finally:
            if saved is not None:
                self._tables, self._table_ids, self._placements, self.coeffs = saved
            else:
                for a in ("_tables", "_table_ids", "_placements"):
                    self.__dict__.pop(a, None)
        self.Hlocal = QMB_hamiltonian(self.lvals, self.sector_dim)
        self.Hlocal.set_matrix(csr, "sparse")
        mp = getattr(self, "momentum_pair", None)
        if mp is not None:
            # the rectangular block P_kL^dagger H_R0 P_kR, as the reference's
            # pair-mode terms give it (host product of the device matrices)
            A = (mp["L"].to_scipy().conj().T @ csr.to_scipy() @ mp["R"].to_scipy()).tocsr()
            A.sum_duplicates()
            self.Hlocal.Ham, self.Hlocal.shape = A, A.shape
            self.Hlocal.nnz = A.nnz
            return A
        return csr

    def get_mask(self, sites_list):
        """Boolean mask over the lattice selecting the given sites
        (SU2_model.py:701-706)."""
        mask = np.zeros(self.lvals, dtype=bool)
        for site in sites_list:
            mask[site] = True
        return mask

    def build_base_Hamiltonian(self, g, m=None, theta=0.0, lambda_noise=0.0, dtype_mode="auto"):
        """The hard-coded spin-1/2 builder by name (SU2_model.py:200-355), for
        models built with that operator set; build_Hamiltonian dispatches to it."""
        if self.use_generic_model:
            raise ValueError("this model carries the generalised operators: use build_gen_Hamiltonian")
        return self.build_Hamiltonian(g, m, theta, lambda_noise, dtype_mode)

    def build_gen_Hamiltonian(self, g, m=None, dtype_mode="auto"):
        """The generalised builder by name (SU2_model.py:357-484), for models
        built with use_generic_model=True or spin >= 1."""
        if not self.use_generic_model:
            raise ValueError("this model carries the hard-coded spin-1/2 operators: construct it with "
                             "use_generic_model=True to build the generalised Hamiltonian")
        return self.build_Hamiltonian(g, m, dtype_mode=dtype_mode)

    # corner combinations of the generalised plaquette (SU2_model.py:420-437)
    _GEN_PLAQ = ["ABABABAB", "AAABBBAB", "ABABAABB", "AAABBABB", "ABBBABAA", "AABBBBAA", "ABBBAABA",
                 "AABBBABA", "BBAAABAB", "BAAABBAB", "BBAAAABB", "BAAABABB", "BBBAABAA", "BABABBAA",
                 "BBBAAABA", "BABABABA"]

    def _build_generic(self, m, real):
        """build_gen_Hamiltonian (SU2_model.py:357-484): electric Casimir,
        staggered mass N_tot, the two hoppings [Q1_p^dag, Q2_m] and
        [Q2_p^dag, Q1_m] per direction and parity, and per plane the sixteen
        species combinations of the plaquette."""
        c = self.coeffs
        lat = self.lattice
        self._new_program()
        self._add_local("E_square", c["E"])
        if not self.pure_theory:
            if m is None:
                raise ValueError("m is required with matter fields")
            even = lat.even_mask()
            self._add_local("N_tot", c["m_even"], [i for i in range(self.n_sites) if even[i]])
            self._add_local("N_tot", c["m_odd"], [i for i in range(self.n_sites) if not even[i]])
            for d in self.directions:
                for stag in ("even", "odd"):
                    bonds = [b for b in lat.bonds(d) if even[b[0]] == (stag == "even")]
                    for names in ([f"Q1_p{d}_dag", f"Q2_m{d}"], [f"Q2_p{d}_dag", f"Q1_m{d}"]):
                        self._add_nbody(names, bonds, c[f"t{d}_{stag}"], True)
        if self.dim > 1:
            planes = ["xy"] if self.dim == 2 else ["xy", "xz", "yz"]
            for a, b in planes:
                plaq = lat.plaquettes((a, b))
                for sp4 in self._GEN_PLAQ:
                    k = [sp4[2 * q:2 * q + 2] for q in range(4)]
                    names = [f"C{k[0]}_p{a},p{b}", f"C{k[1]}_p{b},m{a}", f"C{k[2]}_m{b},p{a}", f"C{k[3]}_m{a},m{b}"]
                    self._add_nbody(names, plaq, c["B"], True)
        return self._build_program(real)

    def local_parity_labels(self, wrt_site):
        """Local action of the spatial inversion on the six dressed states of
        the 1D spin-1/2 chain with matter (SU2_model.py:708-741): (loc_perm,
        loc_phase); wrt_site=False composes it with the particle-hole map."""
        if self.dim > 1:
            raise ValueError(f"Does not work in D={self.dim}>1")
        if self.spin > 0.5:
            raise ValueError(f"Does not work spin j={self.spin}>0.5")
        if wrt_site:
            return np.array([0, 1, 3, 2, 4, 5], dtype=np.int32), np.array([1, -1, 1, -1, 1, -1], dtype=np.int32)
        return np.array([4, 5, 3, 2, 0, 1], dtype=np.int32), np.array([1, -1, 1, 1, 1, -1], dtype=np.int32)

    def get_parity_inversion_operator(self, wrt_site):
        """Spatial inversion of the chain about site L/2 - 1 (wrt_site) or
        about the bond (L/2 - 1, L/2) composed with particle-hole, as a signed
        permutation of the sector states: self.parityOP (SciPy CSR, one +-1
        per column; SU2_model.py:743-763, inversion_sym.py:18-161).  The image
        of every state is found on the device (edlgt_basis_signed_image)."""
        import scipy.sparse as sps
        if self.dim != 1:
            raise NotImplementedError(f"Cannot apply PARITY in D={self.dim}>1")
        loc_perm, loc_phase = self.local_parity_labels(wrt_site)
        L_ = self.n_sites
        center2 = 2 * (L_ // 2 - 1) + (0 if wrt_site else 1)
        site_perm = np.array([(center2 - j) % L_ for j in range(L_)], dtype=np.int32)
        row, sign = self.basis.signed_image(site_perm, loc_perm, loc_phase)
        N = self.sector_dim
        self._parity_image = row
        self.parityOP = sps.csr_matrix((sign, (row, np.arange(N))), shape=(N, N))
        return self.parityOP

    def check_parity_inversion_operator(self):
        """The checks of SU2_model.py:765-829: |PH - HP|, |P^2 - I|, |P^H - P|
        and |P^H P - I| (returned), one +-1 per row and column, and, for every
        basis state i, <i|H|i> = <Pi|H|Pi> and HP|i> = PH|i> (the reference's
        per-state loop, vectorised: the diagonal of H against its permuted
        diagonal, and the column norms of HP - PH) to 1e-14 times the largest
        |H| entry (the reference's 1e-14 taken relative, as the entries carry
        their own rounding); ValueError on the first state that fails.  H is
        taken from the device (host copy)."""
        import scipy.sparse as sps
        from scipy.sparse.linalg import norm as spnorm
        P = self.parityOP
        H = self.H.Ham.to_scipy()
        if self.momentum_basis is not None:
            Pk = self.momentum_basis.to_scipy()
            H = (Pk @ H @ Pk.conj().T).tocsr()
        I = sps.identity(self.sector_dim, format="csr")
        comm = (P @ H - H @ P).tocsc()
        errs = {"PH-HP": float(spnorm(comm)), "P^2-I": float(spnorm(P @ P - I)),
                "P^H-P": float(spnorm(P.conj().T - P)), "P^HP-I": float(spnorm(P.conj().T @ P - I))}
        for k, v in errs.items():
            logger.info(f"|{k}| = {v}")
        if not (np.all(np.diff(P.indptr) == 1) and np.all(np.diff(P.tocsc().indptr) == 1)):
            raise ValueError("the parity operator is not a signed permutation")
        if self.momentum_basis is None:
            d = np.real(H.diagonal())
            img = self._parity_image  # P|i> = +-|img[i]>
            cols = np.sqrt(np.asarray(abs(comm).power(2).sum(axis=0))).ravel()
            tol = 1e-14 * max(1.0, float(abs(H).max()))
            bad = np.nonzero((np.abs(d[img] - d) > tol) | (cols > tol))[0]
            if bad.size:
                ii = int(bad[0])
                raise ValueError(f"config {ii} {self.sector_configs[ii]} not symmetrized")
        return errs

    def overlap_QMB_state(self, name):
        """Reference product configurations (SU2_model.py:486-553), local-state
        indices of the hard-coded spin-1/2 basis."""
        if self.use_generic_model:
            raise NotImplementedError("named states index the hard-coded spin-1/2 basis")
        if self.dim == 1:
            table = {"V": (0, 4, 0, 2), "PV": (1, 5, 1, 1), "T": (6, 12, 1, 1), "M": (2, 3, 1, 1),
                     "B": (7, 11, 1, 1)}
            s1, s2, Lc, Rc = table.get(name, (0, 0, 0, 0))
            cfg = [s1 if i % 2 == 0 else s2 for i in range(self.n_sites)]
            if name == "DW":
                s1, s2, Lc, Rc = 0, 4, 0, 2
                cfg = [s1 if i < self.n_sites // 2 else s2 for i in range(self.n_sites)]
            elif name == "DW2":
                s1, s2, Lc, Rc = 0, 4, 0, 2
                cfg = [s1 if (i // 2) % 2 == 0 else s2 for i in range(self.n_sites)]
            if self.has_obc[0]:
                cfg[0], cfg[-1] = Lc, Rc
            return np.array(cfg)
        if not self.pure_theory:
            if self.has_obc[0]:
                table = {"V": [0, 9, 0, 4, 4, 0, 9, 0], "PV1": [1, 12, 3, 5, 5, 2, 11, 1],
                         "PV2": [1, 11, 1, 5, 5, 3, 10, 1]}
                return np.array(table[name])
            return np.array([0, 9, 0, 9, 9, 0, 9, 0])
        return np.zeros(self.n_sites, dtype=int)
