"""GPU: SU2_Model.get_parity_inversion_operator / check_parity_inversion_operator
(SU2_model.py:708-829): the device images of every sector state against the
oracle's inversion (oracle/inversion.py, pinned in tests/test_inversion.py),
bit-exact; the model's checks pass on periodic chains in both modes and in a
momentum sector; the ground state's parity <psi|P|psi> is +-1; an open chain
(edge sites with three of the six dressed states) raises."""
import numpy as np
import pytest

from oracle.inversion import parity_operator
from oracle.model import SU2Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L", [6, 8, 10])
@pytest.mark.parametrize("wrt_site", [True, False])
def test_parity_operator_matches_the_oracle(gpu, L, wrt_site):
    from edlgt_amd import SU2_Model
    print(f"parity: SU(2) L={L} PBC wrt_site={wrt_site}", flush=True)
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[L], lvals=[L], has_obc=[False])
    orc = SU2Oracle(False, [L], [False], sectors=[L])
    P = mod.get_parity_inversion_operator(wrt_site)
    Po = parity_operator(orc.configs, orc.keys, orc.loc_dims, wrt_site)
    assert np.array_equal(P.indptr, Po.indptr) and np.array_equal(P.indices, Po.indices)
    assert np.array_equal(P.data, Po.data)
    mod.build_Hamiltonian(1.0, 5.0)
    errs = mod.check_parity_inversion_operator()
    assert errs["P^2-I"] == 0.0 and errs["P^H-P"] == 0.0 and errs["P^HP-I"] == 0.0
    assert errs["PH-HP"] < 1e-12
    mod.diagonalize_Hamiltonian(2, "sparse")
    if mod.H.energies[1] - mod.H.energies[0] > 1e-8:
        psi = mod.H.Npsi[0].psi
        assert abs(abs(np.real(np.vdot(psi, P @ psi))) - 1.0) < 1e-9


def test_parity_check_in_a_momentum_sector(gpu):
    """With a momentum basis the check compares P with B H_k B^H (the
    reference's branch, SU2_model.py:767-773)."""
    from edlgt_amd import SU2_Model
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    mod.get_parity_inversion_operator(True)
    mod.set_momentum_sector([2], [0])
    mod.build_Hamiltonian(1.0, 5.0)
    errs = mod.check_parity_inversion_operator()
    assert errs["PH-HP"] < 1e-12


def test_open_chain_raises(gpu):
    from edlgt_amd import SU2_Model
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[True])
    # edge sites hold 3 of the 6 dressed states: the local map leaves their basis
    with pytest.raises(ValueError, match="outside its image site's basis|not closed"):
        mod.get_parity_inversion_operator(True)


def test_parity_penalty_added_to_the_hamiltonian(gpu):
    """The reference's scripts add a parity penalty to H.Ham
    (scripts/script_SU2_dynamics.py: H + 300 (1 - P)/2): the sum of the
    device matrix and SciPy matrices is a device matrix whose spectrum keeps
    the P = +1 levels and lifts the P = -1 ones by 300."""
    import scipy.sparse as sps
    from edlgt_amd import SU2_Model
    from edlgt_amd.device import CSRMatrix
    print("parity penalty", flush=True)
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    mod.build_Hamiltonian(1.0, 5.0)
    mod.get_parity_inversion_operator(wrt_site=False)
    P = mod.parityOP
    Pminus = 0.5 * (sps.identity(P.shape[0], format="csr") - P)
    H0 = mod.H.Ham.to_scipy().toarray()
    before = mod.H.full_eigenpairs()[0]  # cached levels of H, dropped by the assignment below
    mod.H.Ham = mod.H.Ham + 300 * Pminus
    assert isinstance(mod.H.Ham, CSRMatrix)
    Hp = mod.H.Ham.to_scipy().toarray()
    assert np.max(np.abs(Hp - (H0 + 300 * Pminus.toarray()))) < 1e-12
    Pd = P.toarray()
    w, V = np.linalg.eigh(H0 + 1e-9 * Pd)  # levels labelled by parity
    par = np.real(np.einsum("ij,ik,kj->j", V.conj(), Pd, V))
    plus = np.sort(w[par > 0])
    mod.diagonalize_Hamiltonian(3, "sparse", tol=1e-12)
    assert np.max(np.abs(mod.H.energies - plus[:3])) < 1e-6
    assert np.max(np.abs((2.0 * mod.H.Ham).to_scipy().toarray() - 2 * Hp)) == 0.0
    assert np.max(np.abs(before - np.linalg.eigvalsh(H0))) < 1e-9
    assert np.max(np.abs(mod.H.full_eigenpairs()[0] - np.linalg.eigvalsh(Hp))) < 1e-9
