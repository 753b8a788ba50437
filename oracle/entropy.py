"""Bipartite entanglement of a sector state (oracle; test infrastructure).

Restatement of:
  * QuantumModel._get_partition              models/quantum_model.py:715-791
  * QMB_state._get_psi_matrix                modeling/qmb_state.py:160-209
  * QMB_state.reduced_density_matrix         modeling/qmb_state.py:211-243
  * QMB_state.entanglement_entropy           modeling/qmb_state.py:245-291

``dense_*`` follow the reference directly: psi reshaped into the dense
psi[subsystem config, environment config] matrix, one SVD (or M M^H).
``block_schmidt_weights`` computes the same spectrum block by block over the
connected components of the (subsystem, environment) pairs present in the
sector, for sizes where the dense matrix is too large.  Pins: the SU(2)
known-answer entropy (validation/su2/test_su2_spectrum.py, 0.802957045143)
and the equality of the block and dense spectra on random states
(tests/test_su2.py).
"""
import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components


def partition_maps(configs, keep):
    """Subsystem / environment configuration indices of every basis state
    (rows of np.unique over the kept / traced-out site columns)."""
    env = [i for i in range(configs.shape[1]) if i not in keep]
    _, a = np.unique(configs[:, list(keep)], axis=0, return_inverse=True)
    _, e = np.unique(configs[:, env], axis=0, return_inverse=True)
    a, e = a.ravel(), e.ravel()
    return dict(subsys_map=a, env_map=e, subsys_dim=int(a.max()) + 1, env_dim=int(e.max()) + 1)


def dense_psi_matrix(psi, part):
    M = np.zeros((part["subsys_dim"], part["env_dim"]), dtype=np.complex128)
    M[part["subsys_map"], part["env_map"]] = psi
    return M


def dense_schmidt_weights(psi, part):
    return np.linalg.svd(dense_psi_matrix(psi, part), compute_uv=False) ** 2


def dense_rdm(psi, part):
    M = dense_psi_matrix(psi, part)
    return M @ M.conj().T


def entropy(lam):
    """-sum lam log2 lam over lam > 1e-10 (qmb_state.py:280-285)."""
    lam = np.asarray(lam)
    lam = lam[lam > 1e-10]
    return float(-np.sum(lam * np.log2(lam)))


def block_schmidt_weights(psi, part):
    """Squared singular values of psi reshaped by the partition maps, block by
    block over the connected components of the bipartite graph of
    (subsystem, environment) pairs."""
    a, e = np.asarray(part["subsys_map"]), np.asarray(part["env_map"])
    na, ne = part["subsys_dim"], part["env_dim"]
    g = sp.coo_matrix((np.ones(a.size), (a, na + e)), shape=(na + ne, na + ne))
    nc, lab = connected_components(g, directed=False)
    comp = lab[a]
    order = np.argsort(comp, kind="stable")
    bounds = np.searchsorted(comp[order], np.arange(nc + 1))
    out = []
    for c in range(nc):
        idx = order[bounds[c]:bounds[c + 1]]
        if idx.size == 0:
            continue
        ua, ia = np.unique(a[idx], return_inverse=True)
        ue, ie = np.unique(e[idx], return_inverse=True)
        M = np.zeros((ua.size, ue.size), dtype=np.complex128)
        M[ia.ravel(), ie.ravel()] = psi[idx]
        out.append(np.linalg.svd(M, compute_uv=False) ** 2)
    return np.concatenate(out) if out else np.zeros(0)
