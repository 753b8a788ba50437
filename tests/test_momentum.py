"""Momentum sectors: the oracle's projector (a restatement of
edlgt/symmetries/translational_sym.py) pinned by what the mathematics fixes:
orthonormal columns, completeness (the non-empty k-sector dimensions add up
to the sector), H B_k = B_k (B_k^H H B_k) for a translation-invariant H, and the union
of the k-sector spectra equal to the full spectrum."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle.model import SU2Oracle, ZnOracle
from oracle.momentum import momentum_basis, source_sites


def _check_sector_family(configs, keys, loc_dims, lvals, uc, H, ks, dense_spectrum):
    N = configs.shape[0]
    tot, allk = 0, []
    for k in ks:
        B = momentum_basis(configs, keys, loc_dims, lvals, uc, k)
        # a reference whose Bloch sum cancels (a stabilizer off the axes with a
        # non-trivial phase) keeps an empty column, as in the reference
        live = np.diff(B.indptr) > 0
        assert live.all() or any(kk != 0 for kk in k)
        B = B[:, np.nonzero(live)[0]]
        G = (B.conj().T @ B).toarray()
        assert np.max(np.abs(G - np.eye(B.shape[1]))) < 1e-14
        Hk = (B.conj().T @ H @ B).tocsr()
        assert abs(Hk - Hk.conj().T).max() < 1e-13 if Hk.nnz else True
        D = (H @ B - B @ Hk)
        assert (abs(D).max() if D.nnz else 0.0) < 1e-12
        tot += B.shape[1]
        if dense_spectrum:
            allk.append(np.linalg.eigvalsh(Hk.toarray()))
    assert tot == N
    if dense_spectrum:
        ev = np.linalg.eigvalsh(H.toarray())
        assert np.max(np.abs(np.sort(np.concatenate(allk)) - ev)) < 1e-11


def test_source_sites_are_translations():
    src, R = source_sites([6], [2])
    assert list(R) == [3]
    assert list(src[1]) == [4, 5, 0, 1, 2, 3]  # dest d receives site d - 2
    src, R = source_sites([3, 3], [1, 1])
    for s in range(9):
        assert sorted(src[s]) == list(range(9))


def test_su2_chain_momentum_spectra():
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    H = orc.hamiltonian(1.0, 5.0)
    _check_sector_family(orc.configs, orc.keys, orc.loc_dims, [6], [2], H, [[0], [1], [2]], True)


def test_z3_square_momentum_sectors():
    z = ZnOracle(3, [3, 3], [False, False], flux_sector=(0, 0))
    H = z.hamiltonian(1.0, 0.7)
    ks = [[a, b] for a in range(3) for b in range(3)]
    _check_sector_family(z.configs, z.keys, z.loc_dims, [3, 3], [1, 1], sp.csr_matrix(H), ks, False)


def test_zero_k_is_real_and_uniform():
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    B = momentum_basis(orc.configs, orc.keys, orc.loc_dims, [6], [2], [0])
    assert np.max(np.abs(B.data.imag)) == 0.0
    for c in range(B.shape[1]):
        v = B[:, c].data
        assert np.allclose(v, v[0]) and abs(np.sum(np.abs(v) ** 2) - 1) < 1e-14


def test_su2_tc_momentum_sectors():
    """The translation-combined symmetry of the SU(2) chain (shift by one
    site after the local map C): the L sectors k = 0..5 of the L=6 ring
    partition the sector, commute with H and reproduce its spectrum."""
    from oracle.momentum import momentum_basis_tc
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    H = orc.hamiltonian(1.0, 5.0)
    ev = np.linalg.eigvalsh(H.toarray())
    tot, allk = 0, []
    for k in range(6):
        B = momentum_basis_tc(orc.configs, orc.keys, orc.loc_dims, k)
        B = B[:, np.nonzero(np.diff(B.indptr) > 0)[0]]
        G = (B.conj().T @ B).toarray()
        assert np.max(np.abs(G - np.eye(B.shape[1]))) < 1e-14
        Hk = (B.conj().T @ H @ B).tocsr()
        D = H @ B - B @ Hk
        assert (abs(D).max() if D.nnz else 0.0) < 1e-12
        tot += B.shape[1]
        allk.append(np.linalg.eigvalsh(Hk.toarray()))
    assert tot == orc.N
    assert np.max(np.abs(np.sort(np.concatenate(allk)) - ev)) < 1e-11
