"""Ensemble averages over a fully diagonalised sector (oracle restatement;
test infrastructure).

Follows edlgt/models/quantum_model.py: microcanonical_avg (:942-1045; shell
|e_i - <H>/L| < sqrt(<H^2> - <H>^2)/L over per-site energies, equal weights,
coherent shell state), diagonal_avg (:1047-1156; weights |<E_i|psi>|^2, or per
degenerate block (build_energy_block_ids, :1451-1480) the block projection
sum_b <Psi_b|O|Psi_b>), canonical_avg (:912-940; Tr(O e^{-beta H}) / (Z L)
with O the sum of the site operators).  Observables are diagonal local
operators (per-site diagonals `diag`, shape (L, d)); `post` is the reference's
per-observable reduction of the per-site values (mean, staggered mean, or dot
with a normalisation vector / L).  Everything in loops over eigenstates, as
the reference does it.  Beside them the spectral sums behind the reference's
partition_function / thermal_average / F_prime / get_beta / get_r_value
(edlgt/modeling/qmb_hamiltonian.py:428-601), with the dense exponential where
the reference uses it.
"""
import numpy as np


def site_values(psi, configs, diag):
    """<psi|O_s|psi> for every site s of a diagonal local operator."""
    p = np.abs(psi) ** 2
    return np.array([np.sum(p * diag[s][configs[:, s]]) for s in range(configs.shape[1])])


def energy_blocks(evals, tol=1e-10):
    block, b = np.empty(len(evals), dtype=np.int64), 0
    block[0] = 0
    for k in range(1, len(evals)):
        if abs(evals[k] - evals[k - 1]) >= tol:
            b += 1
        block[k] = b
    return block, b + 1


def microcanonical(H, evecs, n_energies, state, configs, diags, post, n_sites):
    Eq = np.vdot(state, H @ state).real
    E2q = np.vdot(H @ state, H @ state).real
    e_q, d_e = Eq / n_sites, np.sqrt(E2q - Eq ** 2) / n_sites
    shell = np.nonzero(np.abs(n_energies - e_q) < d_e)[0]
    if shell.size == 0:
        raise ValueError("empty shell")
    psi_th = np.zeros(H.shape[0], dtype=complex)
    out = {k: 0.0 for k in diags}
    for i in shell:
        psi_th += evecs[:, i]
        for k in diags:
            out[k] += post[k](site_values(evecs[:, i], configs, diags[k]))
    return psi_th / np.sqrt(shell.size), {f"ME_{k}": v / shell.size for k, v in out.items()}


def diagonal(evecs, n_energies, state, configs, diags, post, tol=1e-10):
    block, nb = energy_blocks(n_energies, tol)
    coef = evecs.conj().T @ state
    out = {f"DE_{k}": 0.0 for k in diags}
    if np.all(np.bincount(block) == 1):
        for i in range(evecs.shape[1]):
            p = abs(coef[i]) ** 2
            for k in diags:
                out[f"DE_{k}"] += p * post[k](site_values(evecs[:, i], configs, diags[k]))
    else:
        for b in range(nb):
            idx = np.nonzero(block == b)[0]
            psi_b = evecs[:, idx] @ coef[idx]
            if np.vdot(psi_b, psi_b).real < 1e-14:
                continue
            for k in diags:
                out[f"DE_{k}"] += post[k](site_values(psi_b, configs, diags[k]))
    return out


def canonical(energies, evecs, configs, diag, beta, n_sites):
    w = np.exp(-beta * (energies - energies.min()))
    Z = np.sum(w)
    return float(sum(w[i] * np.sum(site_values(evecs[:, i], configs, diag)) for i in range(len(w))) / (Z * n_sites))


# -- spectral sums and the beta search (qmb_hamiltonian.py:428-601)
def expm_sums(H, beta):
    """(Z, <H>, <H^2>) from the dense exponential, as the reference's branch
    without a stored spectrum computes them (qmb_hamiltonian.py:448-449,
    :487-488, :505-506)."""
    from scipy.linalg import expm
    X = expm(-beta * H)
    Z = np.real(np.trace(X))
    return Z, np.real(np.trace(H @ X)) / Z, np.real(np.trace(H @ H @ X)) / Z


def level_sums(levels, beta):
    """(Z, <e>, <e^2>) as sums over levels (the full-spectrum branch, over the
    per-site energies)."""
    w = np.exp(-beta * levels)
    Z = np.sum(w)
    return Z, levels.dot(w) / Z, np.square(levels).dot(w) / Z


def newton_beta(sums, target, threshold=1e-10, max_iter=1000):
    """The reference's iteration (qmb_hamiltonian.py:531-562): beta from 1e-7,
    beta <- beta - F / F', F = <e> - target, F' = -<e^2> - <e>^2."""
    beta, acc, it = 1e-7, 1.0, 0
    while acc > threshold and it < max_iter:
        it += 1
        _, E, E2 = sums(beta)
        prev, beta = beta, beta - (E - target) / (-E2 - E ** 2)
        acc = abs(beta - prev)
    return beta, it


def r_values(n_energies):
    """Adjacent-gap ratios of the sorted spectrum and their mean over the
    middle 80 % (qmb_hamiltonian.py:566-601), in a loop."""
    d = np.diff(np.sort(n_energies))
    r = np.array([min(d[i], d[i + 1]) / max(d[i], d[i + 1]) for i in range(len(d) - 1)])
    return r, np.mean(r[int(0.1 * len(r)):int(0.9 * len(r))])
