"""GPU: momentum sectors against the oracle's restatement of the reference's
projector (same columns, rows and amplitudes), B^H H B against the oracle's
sparse product, eigenvalues per sector against the oracle, and the sector
ground states against the full-sector ground state and the reference's SU(2)
known answers."""
import numpy as np
import pytest

from oracle.model import SU2Oracle, ZnOracle, lowest_eigs
from oracle.momentum import momentum_basis

pytestmark = pytest.mark.gpu


def _compare(mb, Bo, Hk, Ho):
    B = mb.to_scipy()
    assert B.shape == Bo.shape
    assert np.array_equal(B.indptr, Bo.indptr) and np.array_equal(B.indices, Bo.indices)
    assert np.max(np.abs(B.data - Bo.data), initial=0.0) < 1e-15
    ref = (Bo.conj().T @ Ho @ Bo).toarray()
    got = Hk.to_scipy().toarray()
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref), initial=0.0) < 1e-13 * max(1.0, np.max(np.abs(ref), initial=0.0))


@pytest.mark.parametrize("k", [0, 1, 2])
def test_su2_chain_momentum(gpu, k):
    from edlgt_amd import SU2_Model
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    mod.set_momentum_sector([2], [k])
    Hk = mod.build_Hamiltonian(1.0, 5.0)
    Bo = momentum_basis(orc.configs, orc.keys, orc.loc_dims, [6], [2], [k])
    Ho = orc.hamiltonian(1.0, 5.0)
    _compare(mod.momentum_basis, Bo, Hk, Ho)
    mod.diagonalize_Hamiltonian(3, "sparse")
    ref = np.linalg.eigvalsh((Bo.conj().T @ Ho @ Bo).toarray())[:3]
    assert np.max(np.abs(mod.H.energies - ref)) < 1e-10


def test_su2_ground_state_in_k0_matches_kat(gpu):
    """The L=6 PBC ground state lies in k = 0: energy and observables of the
    reference's known answers (validation/su2/test_su2_spectrum.py)."""
    from edlgt_amd import SU2_Model
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    mod.set_momentum_sector([2], [0])
    mod.build_Hamiltonian(1.0, 5.0)
    mod.diagonalize_Hamiltonian(1, "sparse")
    assert abs(mod.H.Nenergies[0] - (-6.173752796132477)) < 1e-10
    mod.get_observables(["T2_px", "N_single", "N_pair", "N_zero"])
    mod.measure_observables(0)
    e2 = mod.link_avg("T2")
    assert abs(e2 - 0.0553604161539725) < 1e-10
    assert abs(mod.stag_avg("N_single") - 0.14157936434695081) < 1e-10


def test_z3_square_momentum_sectors(gpu):
    from edlgt_amd import Zn_Model
    z = ZnOracle(3, [3, 3], [False, False], flux_sector=(0, 0))
    Ho = z.hamiltonian(1.0, 0.7)
    e_full = lowest_eigs(Ho, 1)[0]
    best = np.inf
    tot = 0
    for kx in range(3):
        for ky in range(3):
            mod = Zn_Model(3, lvals=[3, 3], has_obc=[False, False], flux_sector=(0, 0))
            mod.set_momentum_sector([1, 1], [kx, ky])
            Hk = mod.build_Hamiltonian(1.0, 0.7)
            Bo = momentum_basis(z.configs, z.keys, z.loc_dims, [3, 3], [1, 1], [kx, ky])
            _compare(mod.momentum_basis, Bo, Hk, Ho)
            tot += int(np.sum(np.diff(Bo.indptr) > 0))
            mod.diagonalize_Hamiltonian(1, "sparse")
            best = min(best, mod.H.energies[0])
    assert tot == z.N
    assert abs(best - e_full) < 1e-10


def test_momentum_errors(gpu):
    from edlgt_amd import SU2_Model
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[True])
    with pytest.raises(ValueError):
        mod.set_momentum_sector([2], [0])
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    with pytest.raises(ValueError):
        mod.set_momentum_sector([4], [0])  # 4 does not divide 6
    with pytest.raises(ValueError):
        mod.set_momentum_sector([2], [0, 1])


@pytest.mark.parametrize("k", range(6))
def test_su2_tc_momentum(gpu, k):
    """TC_symmetry=True (unit cell 1, the L momenta of X = T o C) against the
    oracle's restatement; the lowest sector eigenvalue over all k is the
    reference's known answer."""
    from edlgt_amd import SU2_Model
    from oracle.momentum import momentum_basis_tc
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    mod.set_momentum_sector([1], [k], TC_symmetry=True)
    Hk = mod.build_Hamiltonian(1.0, 5.0)
    Bo = momentum_basis_tc(orc.configs, orc.keys, orc.loc_dims, k)
    _compare(mod.momentum_basis, Bo, Hk, orc.hamiltonian(1.0, 5.0))
    mod.diagonalize_Hamiltonian(1, "sparse")
    if k == 0:
        assert abs(mod.H.Nenergies[0] - (-6.173752796132477)) < 1e-10
    else:
        assert mod.H.Nenergies[0] > -6.173752796132477 - 1e-10


def _finite_k_par(model, k_vals, unit_cell, **couplings):
    par = {"model": model, "hamiltonian": {"n_eigs": 1, "save_psi": False},
           "momentum": {"get_momentum_basis": True, "unit_cell_size": unit_cell, "momentum_k_vals": k_vals,
                        "TC_symmetry": False},
           "observables": {"measure_obs": False, "get_entropy": False, "get_state_configs": False,
                           "get_overlap": False}}
    par.update(couplings)
    return par


def test_reference_finite_k_models(gpu):
    """The two models of the reference's finite-momentum check
    (validation/symmetries/test_momentum_nbody_generic_finite_k_models.py:40-136),
    run through the parameter-dictionary drivers with its parameters: the
    6-site periodic SU(2) chain with matter at unit cell 1 and k = pi (g = 1,
    m = 5), and the 2x2x2 periodic pure QED lattice with theta = 0.41 at
    k = (pi, pi, pi) (g = 2.3; complex matrix, 146,343-state sector).  The
    device projector and the ground-state energy are compared with the oracle's
    projector and the lowest eigenvalue of its B^H H B."""
    from edlgt_amd import QED_Model, SU2_Model
    from edlgt_amd.workflows import run_QED_spectrum, run_SU2_spectrum
    from oracle.model import QEDOracle
    print("finite k: SU(2) L=6 PBC, unit cell 1, k = 3", flush=True)
    par = _finite_k_par({"lvals": [6], "sectors": [6], "has_obc": [False], "spin": 0.5, "pure_theory": False,
                         "ham_format": "sparse"}, [3], [1], g=1.0, m=5.0)
    res = run_SU2_spectrum(par)
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    Bo = momentum_basis(orc.configs, orc.keys, orc.loc_dims, [6], [1], [3])
    ref = np.linalg.eigvalsh((Bo.conj().T @ orc.hamiltonian(1.0, 5.0) @ Bo).toarray())[0] / 6
    assert abs(float(res["energy"][0]) - ref) < 1e-10, (res["energy"][0], ref)
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    mod.set_momentum_sector([1], [3])
    _compare(mod.momentum_basis, Bo, mod.build_Hamiltonian(1.0, 5.0), orc.hamiltonian(1.0, 5.0))

    print("finite k: QED 2x2x2 PBC pure, theta = 0.41, k = (1, 1, 1)", flush=True)
    par = _finite_k_par({"lvals": [2, 2, 2], "has_obc": [False] * 3, "spin": 1, "pure_theory": True,
                         "ham_format": "sparse"}, [1, 1, 1], [1, 1, 1], g=2.3, theta=0.41)
    res = run_QED_spectrum(par)
    q = QEDOracle(1, True, [2, 2, 2], [False] * 3)
    Hq = q.hamiltonian(2.3, None, 0.41)
    Bq = momentum_basis(q.configs, q.keys, q.loc_dims, [2, 2, 2], [1, 1, 1], [1, 1, 1])
    ref = lowest_eigs((Bq.conj().T @ Hq @ Bq).tocsr(), 1)[0] / 8
    assert abs(float(res["energy"][0]) - ref) < 1e-10, (res["energy"][0], ref)
    mod = QED_Model(spin=1, pure_theory=True, lvals=[2, 2, 2], has_obc=[False] * 3)
    mod.set_momentum_sector([1, 1, 1], [1, 1, 1])
    Hk = mod.build_Hamiltonian(2.3, None, 0.41)
    B = mod.momentum_basis.to_scipy()
    assert B.shape == Bq.shape and np.array_equal(B.indptr, Bq.indptr) and np.array_equal(B.indices, Bq.indices)
    assert np.max(np.abs(B.data - Bq.data)) < 1e-15
    diff = (Hk.to_scipy() - (Bq.conj().T @ Hq @ Bq)).tocsr()  # 18,255 momentum states: compared sparse
    assert np.max(np.abs(diff.data), initial=0.0) < 1e-13 * np.max(np.abs(Hq.data))


@pytest.mark.parametrize("lvals,sites2,sites4", [([6], [1, 4], [0, 2, 3, 5]), ([2, 1, 2], [0, 3], [0, 1, 2, 3])])
def test_reference_generic_nbody_momentum(gpu, lvals, sites2, sites4):
    """The reference's generic n-body momentum cases
    (validation/symmetries/test_momentum_nbody_generic.py:16-41, 164-184): the
    full space of qubits on a 6-site ring and on a 2x1x2 lattice, k = 0 of the
    unit translations, and a two-site and a four-site product of its complex
    2x2 operators (op f: [[0.7 c, 0.2 + 0.05i c], [-0.3 + 0.11i c, -0.4 c]],
    c = f + 1) on non-adjacent sites.  B^H O B on the device (a stored CSR of
    O projected by edlgt_csr_momentum) against the oracle's projector applied
    to the Kronecker product of the factors."""
    import scipy.sparse as sp
    import edlgt_amd._lib as L
    from edlgt_amd.device import CSRMatrix, MomentumBasis, SectorBasis
    from oracle.model import encode
    n = int(np.prod(lvals))
    configs = np.array(np.unravel_index(np.arange(2 ** n), [2] * n)).T.astype(np.uint16)
    keys = encode(configs, np.full(n, 2))
    Bo = momentum_basis(configs, keys, np.full(n, 2), lvals, [1] * len(lvals), [0] * len(lvals))
    basis = SectorBasis(np.full(n, 2))
    assert np.array_equal(basis.configs(), configs)
    mb = MomentumBasis(basis, lvals, [1] * len(lvals), [0] * len(lvals))
    B = mb.to_scipy()
    assert np.array_equal(B.indptr, Bo.indptr) and np.array_equal(B.indices, Bo.indices)
    assert np.max(np.abs(B.data - Bo.data)) < 1e-15
    for sites in (sites2, sites4):
        ops = [np.array([[0.7 * (f + 1), 0.2 + 0.05j * (f + 1)], [-0.3 + 0.11j * (f + 1), -0.4 * (f + 1)]])
               for f in range(len(sites))]
        H = CSRMatrix(basis, np.stack(ops), [dict(sites=list(sites), tables=list(range(len(sites))), strength=1.0)],
                      L.VALUE_C128)
        full = sp.identity(1, dtype=complex, format="csr")
        for s in range(n):
            f = sites.index(s) if s in sites else None
            full = sp.kron(full, sp.csr_matrix(ops[f]) if f is not None else sp.identity(2), format="csr")
        want = (Bo.conj().T @ full @ Bo).toarray()
        got = H.momentum(mb).to_scipy().toarray()
        assert got.shape == want.shape
        assert np.max(np.abs(got - want)) < 1e-13


def test_momentum_projector_as_scipy(gpu):
    """_basis_Pk_as_csr / _basis_Pk_as_csc (quantum_model.py:615-632): the
    momentum projector as SciPy, None without a momentum sector."""
    from edlgt_amd import SU2_Model
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    assert mod._basis_Pk_as_csr() is None
    mod.set_momentum_sector([2], [1])
    B = mod.momentum_basis.to_scipy()
    for M in (mod._basis_Pk_as_csr(), mod._basis_Pk_as_csc()):
        assert M.shape == B.shape and abs(M - B).max() == 0
