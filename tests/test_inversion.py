"""CPU: the spatial inversion of the 1D SU(2) chain (oracle/inversion.py,
restating edlgt/symmetries/inversion_sym.py with SU2_Model.local_parity_labels).
Pinned by what the symmetry fixes: P^2 = I, P = P^T (a signed permutation that
is its own inverse), and [P, H] = 0 for the oracle's periodic-chain Hamiltonian
at several couplings, for the site-centred inversion and for the bond-centred
one composed with particle-hole; a planted mistake (the centre moved by one
site, or the two local maps exchanged) breaks the commutation, so the check
has teeth; the non-degenerate ground state is a P eigenvector (eigenvalue +-1).
An open chain is not inversion-symmetric about these centres: an image leaves
the sector."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle.inversion import SU2_PARITY_BOND, SU2_PARITY_SITE, parity_operator
from oracle.model import SU2Oracle, lowest_eigs


def _comm(P, H):
    return float(abs(P @ H - H @ P).max()) / float(abs(H).max())


@pytest.mark.parametrize("L", [6, 8])
@pytest.mark.parametrize("wrt_site", [True, False])
def test_inversion_is_a_symmetry(L, wrt_site):
    orc = SU2Oracle(False, [L], [False], sectors=[L])
    P = parity_operator(orc.configs, orc.keys, orc.loc_dims, wrt_site)
    I = sp.identity(orc.N, format="csr")
    assert abs(P @ P - I).max() == 0 and abs(P - P.T).max() == 0
    assert np.all(np.diff(P.indptr) == 1)
    for g, m in ((1.0, 5.0), (0.4, 0.3)):
        H = orc.hamiltonian(g, m)
        assert _comm(P, H) < 1e-14
    e, v = np.linalg.eigh(orc.hamiltonian(1.0, 5.0).toarray()) if orc.N < 3000 else (None, None)
    if e is not None and e[1] - e[0] > 1e-8:
        p = float(np.real(np.vdot(v[:, 0], P @ v[:, 0])))
        assert abs(abs(p) - 1.0) < 1e-10


@pytest.mark.parametrize("wrt_site", [True, False])
def test_planted_mistakes_break_the_commutation(wrt_site):
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    H = orc.hamiltonian(1.0, 5.0)
    other = SU2_PARITY_BOND if wrt_site else SU2_PARITY_SITE
    P = parity_operator(orc.configs, orc.keys, orc.loc_dims, wrt_site, loc_maps=other)
    assert _comm(P, H) > 1e-3
    # centre moved by one site: the site-centred map on the bond-centred
    # geometry and the other way round
    P = parity_operator(orc.configs, orc.keys, orc.loc_dims, not wrt_site,
                        loc_maps=SU2_PARITY_SITE if wrt_site else SU2_PARITY_BOND)
    assert _comm(P, H) > 1e-3


def test_open_chain_images_leave_the_sector():
    orc = SU2Oracle(False, [6], [True], sectors=[6])
    with pytest.raises(ValueError):
        parity_operator(orc.configs, orc.keys, orc.loc_dims, True)


def test_engine_local_maps_are_the_reference_tables():
    import edlgt_amd.models as models
    m = object.__new__(models.SU2_Model)
    m.dim, m.spin = 1, 0.5
    for ws, (perm, phase) in ((True, SU2_PARITY_SITE), (False, SU2_PARITY_BOND)):
        lp, ph = m.local_parity_labels(ws)
        assert np.array_equal(lp, perm) and np.array_equal(ph, phase)
    m.dim = 2
    with pytest.raises(ValueError):
        m.local_parity_labels(True)
