"""Pins the oracle's sector and triplet builders against brute force:
exhaustive sector enumeration (sitebased_sym_sector_configs,
symmetry_sector.py:27-65) and the full-space Kronecker Hamiltonian restricted
to the sector (validation/qed/test_qed_1d_integrated.py:98-111)."""
import numpy as np
import pytest

from oracle.brute import qed_full_hamiltonian, restrict_to_sector
from oracle.model import QEDOracle, sector_configs_exhaustive


CASES = [([4], [True], False, 1.0, 1.0), ([4], [False], False, 0.8, 0.3), ([6], [True], False, 1.3, 0.7),
         ([3, 2], [True, True], False, 1.1, 0.4), ([2, 2], [True, True], True, 0.9, None)]


@pytest.mark.parametrize("lvals,obc,pure,g,m", CASES)
def test_triplets_equal_bruteforce(lvals, obc, pure, g, m):
    M = QEDOracle(1, pure, lvals, obc)
    H = M.hamiltonian(g, m).toarray()
    Hb = restrict_to_sector(qed_full_hamiltonian(M, g, m), M.configs, M.loc_dims).toarray()
    assert np.max(np.abs(H - Hb)) < 1e-13


@pytest.mark.parametrize("lvals,obc,pure", [([4], [True], False), ([4], [False], False), ([3, 2], [True, True], False)])
def test_sector_equals_exhaustive(lvals, obc, pure):
    M = QEDOracle(1, pure, lvals, obc)
    dg = lambda name: np.real(np.einsum("sii->si", M.ops[name]))
    from oracle import lattice as lat
    pairs = lat.link_pairs(M.lvals, M.has_obc)
    Ep = {d: dg(f"E_p{d}") for d in M.dirs}
    Em = {d: dg(f"E_m{d}") for d in M.dirs}
    N = dg("N")

    def check(c):
        if not pure and abs(sum(N[s, c[s]] for s in range(M.n_sites)) - M.n_sites // 2) > 1e-10:
            return False
        for k, d in enumerate(M.dirs):
            for (i, j) in pairs[k]:
                if abs(Ep[d][i, c[i]] - Em[d][j, c[j]]) > 1e-10:
                    return False
        return True

    ex = sector_configs_exhaustive(M.loc_dims, check)
    assert np.array_equal(ex, M.configs)


def test_hermitian_and_real_when_pure():
    M = QEDOracle(1, True, [3, 3], [False, False])
    H = M.hamiltonian(1.0)
    assert H.dtype == np.float64
    assert abs(H - H.T).max() < 1e-15
