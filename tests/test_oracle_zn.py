"""Pins the Z_n oracle: local basis size 27 for Z_3 (Zn_operators.py:269-360),
corner operators commute (Zn_operators.py:262-265), independent sector counts
(3^(L_x L_y + 1) all-flux, 3^(L_x L_y - 1) per flux sector on PBC lattices),
the J = 0 closed form E_0 = -2 h n_links, and the sector Hamiltonian against the
full-space Kronecker Hamiltonian."""
import numpy as np
import pytest

from oracle import zn_ops
from oracle.brute import restrict_to_sector, zn_full_hamiltonian
from oracle.model import ZnOracle, lowest_eigs


def test_local_basis_and_corners():
    basis, states = zn_ops.gauge_invariant_states(3)
    assert basis["site"].shape == (81, 27)
    ops = zn_ops.dressed_site_operators(3)
    names = ["C_px,py", "C_py,mx", "C_mx,my", "C_my,px"]
    for a in names:
        for b in names:
            assert abs(ops[a] @ ops[b] - ops[b] @ ops[a]).max() == 0


@pytest.mark.parametrize("lv", [[2, 2], [3, 2], [3, 3]])
def test_sector_counts(lv):
    n_sites = lv[0] * lv[1]
    assert ZnOracle(3, lv, [False, False], None).N == 3 ** (n_sites + 1)
    assert ZnOracle(3, lv, [False, False], (0, 0)).N == 3 ** (n_sites - 1)


@pytest.mark.parametrize("lv,fs", [([2, 2], None), ([3, 3], (0, 0)), ([3, 2], (1, 0))])
def test_closed_form_J0(lv, fs):
    M = ZnOracle(3, lv, [False, False], fs)
    e = lowest_eigs(M.hamiltonian(0.0, 0.7), 1)[0]
    if fs in (None, (0, 0)):
        assert abs(e - (-2 * 0.7 * 2 * lv[0] * lv[1])) < 1e-12


@pytest.mark.parametrize("lv,obc,fs", [([2, 2], [False, False], None), ([2, 2], [False, False], (0, 0)),
                                       ([2, 2], [False, False], (1, 2))])
def test_sector_hamiltonian_equals_bruteforce(lv, obc, fs):
    M = ZnOracle(3, lv, obc, fs)
    H = M.hamiltonian(1.3, 0.6).toarray()
    Hb = restrict_to_sector(zn_full_hamiltonian(M, 1.3, 0.6), M.configs, M.loc_dims).toarray()
    assert np.max(np.abs(H - Hb)) < 1e-13
