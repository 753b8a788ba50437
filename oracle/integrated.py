"""1D QED with the gauge field integrated out (oracle; test infrastructure).

Restatement of the direct construction the reference validates its
integrated model against (validation/qed/test_qed_1d_integrated.py:14-68,
operators QED_gauge_integrated_operators, operators/QED_operators.py:541-545
and get_Pauli_operators, operators/spin_operators.py:62-72):

    H = sum_i m (-1)^i N_i
      + sum_{i<L-1} [t S+_i S-_{i+1} + conj(t) S-_i S+_{i+1}],  t = -i/2
      + (g/2) sum_{n<L-1} [sum_{k<=n} (q_k + (Sz_k + (-1)^k)/2)]^2

on the full 2^L space (site 0 most significant, local state 0 = occupied),
then restricted to the sector N_tot = L // 2 in lexicographic order, and the
link fields E_n = sum_{k<=n} (q_k + N_k + ((-1)^k - 1)/2).  Pins (tests):
at g = 0 the lowest eigenvalue is the sum of the L/2 lowest single-particle
energies of the staggered tight-binding chain (Jordan-Wigner, nearest
neighbours, no string), and the electric energy of product states follows
Gauss's law (zero for the bare vacuum, g/2 per unit of flux per link for a
meson or a static charge).
"""
from itertools import product

import numpy as np
import scipy.sparse as sp

I2 = np.eye(2)
SZ = np.diag([1.0, -1.0])
SP = np.array([[0.0, 1.0], [0.0, 0.0]])
SM = SP.T.copy()
NOP = 0.5 * (SZ + I2)


def site_op(op, i, L):
    out = sp.identity(1, format="csr", dtype=np.complex128)
    for s in range(L):
        out = sp.kron(out, sp.csr_matrix(op if s == i else I2), format="csr")
    return out


def sector_configs(L):
    """Configurations (local index per site) with N_tot = L // 2, lexicographic."""
    cfg = np.array(list(product(range(2), repeat=L)), dtype=np.int64)
    occ = (cfg == 0).sum(axis=1)
    return cfg[occ == L // 2]


def full_index(cfg):
    L = cfg.shape[1]
    return (cfg * (2 ** np.arange(L - 1, -1, -1))).sum(axis=1)


def hamiltonian_full(L, g, m, q):
    dim = 2 ** L
    H = sp.csr_matrix((dim, dim), dtype=np.complex128)
    for i in range(L):
        H = H + (m if i % 2 == 0 else -m) * site_op(NOP, i, L)
    t = -0.5j
    for i in range(L - 1):
        H = H + t * (site_op(SP, i, L) @ site_op(SM, i + 1, L))
        H = H + np.conj(t) * (site_op(SM, i, L) @ site_op(SP, i + 1, L))
    ident = sp.identity(dim, format="csr", dtype=np.complex128)
    for n in range(L - 1):
        E = sp.csr_matrix((dim, dim), dtype=np.complex128)
        for k in range(n + 1):
            stag = 1.0 if k % 2 == 0 else -1.0
            E = E + (q[k] + 0.5 * stag) * ident + 0.5 * site_op(SZ, k, L)
        H = H + (g / 2.0) * (E @ E)
    return H


def hamiltonian(L, g, m, q):
    """Sector Hamiltonian (rows/columns in sector_configs order)."""
    idx = full_index(sector_configs(L))
    return hamiltonian_full(L, g, m, q)[idx][:, idx].tocsr()


def link_e2(L, q, psi):
    """<psi| E_n^2 |psi> for n < L-1, psi over the sector."""
    cfg = sector_configs(L)
    occ = (cfg == 0).astype(float)
    out = np.zeros(L - 1)
    for n in range(L - 1):
        off = sum(q[k] + (0.0 if k % 2 == 0 else -1.0) for k in range(n + 1))
        En = off + occ[:, : n + 1].sum(axis=1)
        out[n] = float(np.sum(np.abs(psi) ** 2 * En ** 2))
    return out


def free_fermion_e0(L, m):
    """Ground energy at g = 0: the L/2 lowest eigenvalues of the one-particle
    matrix h_ii = m (-1)^i, h_{i,i+1} = -i/2, h_{i+1,i} = +i/2."""
    h = np.zeros((L, L), dtype=complex)
    for i in range(L):
        h[i, i] = m if i % 2 == 0 else -m
    for i in range(L - 1):
        h[i, i + 1] = -0.5j
        h[i + 1, i] = 0.5j
    return float(np.sum(np.linalg.eigvalsh(h)[: L // 2]))
