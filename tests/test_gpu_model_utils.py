"""GPU: the remaining model utilities of the reference's API on the 6-site
periodic SU(2) chain (282 states).
  * ham_format="dense": time evolution from every eigenpair
    (qmb_hamiltonian.py:363-376, exact_time_evolution :792-836) against the
    oracle's expm_multiply at 1e-9 on a uniform grid and against the dense
    exponential on a non-uniform one (which the Krylov path refuses);
  * _project_state_with_basis (quantum_model.py:634-671): B^H x with the
    oracle's momentum projector B at 1e-14;
  * momentum_basis_projection (quantum_model.py:673-713): "H" built in the
    sector basis, projected on the device in place, equal to the matrix built
    after set_momentum_sector (same pattern, values at 1e-14) and to the
    oracle's B^H H B; a host operator projected by SciPy equals B^H A B."""
import numpy as np
import pytest

from oracle.model import SU2Oracle, expm_path
from oracle.momentum import momentum_basis

pytestmark = pytest.mark.gpu


def _chain(**kw):
    from edlgt_amd import SU2_Model
    return SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False], **kw)


@pytest.fixture(scope="module")
def orc():
    return SU2Oracle(False, [6], [False], sectors=[6])


def test_dense_format_exact_evolution(gpu, orc):
    from scipy.linalg import expm
    print("dense-format exact evolution", flush=True)
    mod = _chain(ham_format="dense")
    mod.build_Hamiltonian(1.0, 5.0)
    psi0 = mod.get_qmb_state_from_configs([mod.overlap_QMB_state("V")])
    psi0 = psi0.cpu().numpy() if hasattr(psi0, "data_ptr") else np.asarray(psi0)
    tl = np.arange(21) * 0.05
    mod.time_evolution_Hamiltonian(psi0, tl)
    assert mod.H.n_eigs == orc.N  # every eigenpair solved first, as the reference does
    H = orc.hamiltonian(1.0, 5.0)
    ref = expm_path(H, np.asarray(psi0, dtype=complex), tl[0], tl[-1], len(tl))
    got = np.stack([s.psi for s in mod.H.psi_time])
    assert np.max(np.abs(got - ref)) < 1e-9
    # any spacing: against the dense exponential
    tl2 = np.array([0.0, 0.013, 0.2, 0.71, 1.9])
    mod.time_evolution_Hamiltonian(psi0, tl2)
    Hd = H.toarray()
    for k, t in enumerate(tl2):
        assert np.max(np.abs(mod.H.psi_time[k].psi - expm(-1j * t * Hd) @ psi0)) < 1e-10
    mod.get_observables(["N_single"])
    mod.measure_observables(4, dynamics=True)
    assert np.all(np.isfinite(mod.res["N_single"]))
    # the sparse format keeps the Krylov path (uniform grids only) ...
    sp = _chain()
    sp.build_Hamiltonian(1.0, 5.0)
    with pytest.raises(ValueError):
        sp.time_evolution_Hamiltonian(psi0, tl2)
    assert sp.H.get_sparsity() == H.nnz / orc.N ** 2
    # ... until converted (qmb_hamiltonian.py:227-258)
    sp.H.convert_hamiltonian("dense")
    sp.time_evolution_Hamiltonian(psi0, tl2)
    assert np.max(np.abs(sp.H.psi_time[-1].psi - expm(-1j * tl2[-1] * Hd) @ psi0)) < 1e-10
    with pytest.raises(ValueError):
        sp.H.convert_hamiltonian("csr")


def test_momentum_state_and_operator_projection(gpu, orc):
    import torch
    print("momentum projections", flush=True)
    Bo = momentum_basis(orc.configs, orc.keys, orc.loc_dims, [6], [2], [1])
    Ho = orc.hamiltonian(1.0, 5.0)
    # H built in the sector basis, then projected in place
    mod = _chain()
    mod.build_Hamiltonian(1.0, 5.0)
    x = np.random.default_rng(3).normal(size=orc.N) + 1j * np.random.default_rng(4).normal(size=orc.N)
    assert np.array_equal(mod._project_state_with_basis(x), x.astype(np.complex128))
    mod.set_momentum_sector([2], [1])
    assert np.max(np.abs(mod._project_state_with_basis(x) - Bo.conj().T @ x)) < 1e-13
    assert np.max(np.abs(mod._project_state_with_basis(torch.as_tensor(x)) - Bo.conj().T @ x)) < 1e-13
    mod.momentum_basis_projection("H")
    assert mod.H.shape == (Bo.shape[1], Bo.shape[1])
    ref = _chain()
    ref.set_momentum_sector([2], [1])
    Hk = ref.build_Hamiltonian(1.0, 5.0)
    a, b = mod.H.Ham.to_scipy(), Hk.to_scipy()
    assert np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
    assert np.max(np.abs(a.data - b.data)) < 1e-14 * np.max(np.abs(b.data))
    want = (Bo.conj().T @ Ho @ Bo).toarray()
    assert np.max(np.abs(a.toarray() - want)) < 1e-13 * np.max(np.abs(want))
    mod.momentum_basis_projection("H")  # already projected: unchanged
    assert mod.H.shape == (Bo.shape[1], Bo.shape[1])
    mod.diagonalize_Hamiltonian(2, "sparse")
    assert np.max(np.abs(mod.H.energies - np.linalg.eigvalsh(want)[:2])) < 1e-10
    # a host operator over the sector basis
    A = np.diag(np.real(np.einsum("sii->si", orc.ops["N_single"]))[np.arange(6)[None, :], orc.configs].sum(axis=1))
    got = mod.momentum_basis_projection(A).toarray()
    assert np.max(np.abs(got - (Bo.conj().T @ A @ Bo))) < 1e-13
    with pytest.raises(ValueError):
        _chain().momentum_basis_projection("H")


def test_projector_from_sector_to_fullspace(gpu, orc):
    """build_projector_from_sector_to_fullspace (quantum_model.py:793-824):
    an isometry whose embedding of the sector reduced density matrix of the
    ground state (sites 0-2 of the 6-site chain) equals the reduced density
    matrix of the same state written out on the full 6^6 product space."""
    print("sector-to-fullspace projector", flush=True)
    mod = _chain()
    mod.build_Hamiltonian(1.0, 5.0)
    mod.diagonalize_Hamiltonian(1, "sparse", tol=1e-12)
    keep = [0, 1, 2]
    part = mod._get_partition(keep)
    P = mod.build_projector_from_sector_to_fullspace(keep)
    d = int(np.prod(mod.loc_dims[keep]))
    assert P.shape == (d, part["subsys_dim"])
    assert np.array_equal((P.T @ P).toarray(), np.eye(part["subsys_dim"]))
    rho = mod.H.Npsi[0].reduced_density_matrix(keep, mod._partition_cache)
    full = (P @ rho @ P.T.toarray())
    dims = np.asarray(mod.loc_dims, dtype=np.int64)
    psi = mod.H.Npsi[0].psi
    cfg = np.asarray(mod.sector_configs, dtype=np.int64)
    sub = cfg[:, :3] @ np.array([dims[1] * dims[2], dims[2], 1])
    env = cfg[:, 3:] @ np.array([dims[4] * dims[5], dims[5], 1])
    M = np.zeros((d, int(np.prod(dims[3:]))), dtype=complex)
    M[sub, env] = psi
    assert np.max(np.abs(full - M @ M.conj().T)) < 1e-13


def test_link_symmetry_checks(gpu):
    """check_symmetries (QED_model.py:707-717, SU2_model.py:474-484) on
    measured fields of ground states: satisfied, and a planted violation in
    the measured values is reported with its site."""
    from edlgt_amd import QED_Model
    print("link symmetry checks", flush=True)
    q = QED_Model(spin=1, pure_theory=False, lvals=[2, 2], has_obc=[False, False])
    q.build_Hamiltonian(1.0, 0.5)
    q.diagonalize_Hamiltonian(1, "sparse")
    q.get_observables([f"E_{s}{d}" for s in "pm" for d in q.directions])
    q.measure_observables(0)
    q.check_symmetries()
    q.res["E_px"] = np.array(q.res["E_px"], dtype=float)
    q.res["E_px"][3] += 1e-6
    with pytest.raises(ValueError, match="x-Link Symmetry is violated at index 3"):
        q.check_symmetries()
    s = _chain()
    s.build_Hamiltonian(1.0, 5.0)
    s.diagonalize_Hamiltonian(1, "sparse")
    s.get_observables(["T2_px", "T2_mx"])
    s.measure_observables(0)
    s.check_symmetries()


def test_state_from_host_array(gpu, orc):
    """QMB_state(psi, lvals, loc_dims) as the reference constructs it
    (qmb_state.py:44-77): a host array is copied to the device; the entropy
    and expectation values through it equal those of the device-built state."""
    print("state from a host array", flush=True)
    from edlgt_amd import QMB_state
    mod = _chain()
    mod.build_Hamiltonian(1.0, 5.0)
    mod.diagonalize_Hamiltonian(1, "sparse", tol=1e-12)
    gs = mod.H.Npsi[0]
    st = QMB_state(psi=gs.psi.copy(), lvals=mod.lvals, loc_dims=mod.loc_dims)
    assert st.psi_device.is_cuda and np.array_equal(st.psi, gs.psi)
    mod._get_partition([0, 1, 2])
    assert abs(st.entanglement_entropy([0, 1, 2], mod._partition_cache)
               - gs.entanglement_entropy([0, 1, 2], mod._partition_cache)) < 1e-14
    assert abs(st.expectation_value(mod.H.Ham) - gs.expectation_value(mod.H.Ham)) < 1e-12
    with pytest.raises(ValueError):
        QMB_state(np.zeros((2, 2)), mod.lvals, mod.loc_dims)


def test_hamiltonian_from_added_terms(gpu, orc):
    """QMB_hamiltonian.add_term / build (qmb_hamiltonian.py:155-225): the
    oracle's SU(2) Hamiltonian handed over in three parts (triplets, a dense
    block, a SciPy matrix) and a zero-sum pair of duplicates is SciPy's
    coalesced matrix on the device (pattern bit-exact, explicit zero kept);
    its lowest eigenvalues match the oracle's; a real matrix is stored as
    float64."""
    import scipy.sparse as sps
    from edlgt_amd import QMB_hamiltonian
    from oracle.model import lowest_eigs
    print("Hamiltonian from added terms", flush=True)
    Ho = orc.hamiltonian(1.0, 5.0).tocoo()
    k = Ho.nnz // 3
    H = QMB_hamiltonian([6], orc.N)
    H.add_term((Ho.row[:k], Ho.col[:k], Ho.data[:k]))
    D = np.zeros((orc.N, orc.N), dtype=complex)
    D[Ho.row[k:2 * k], Ho.col[k:2 * k]] = Ho.data[k:2 * k]
    H.add_term(D)
    H.add_term(sps.csr_matrix((Ho.data[2 * k:], (Ho.row[2 * k:], Ho.col[2 * k:])), shape=Ho.shape))
    H.add_term((np.array([0, 0]), np.array([5, 5]), np.array([0.25, -0.25])))
    assert H.row_list.size == Ho.nnz + 2
    H.build("sparse")
    want = sps.csr_matrix((np.concatenate([Ho.data, [0.25, -0.25]]),
                           (np.concatenate([Ho.row, [0, 0]]), np.concatenate([Ho.col, [5, 5]]))), shape=Ho.shape)
    want.sum_duplicates()
    got = H.Ham.to_scipy()
    assert np.array_equal(got.indptr, want.indptr) and np.array_equal(got.indices, want.indices)
    assert np.max(np.abs(got.data - want.data)) <= 1e-15 * np.max(np.abs(want.data))
    H.diagonalize(3, "sparse", np.full(6, 6))
    assert np.max(np.abs(H.energies - lowest_eigs(Ho.tocsr(), 3))) < 1e-10
    R = QMB_hamiltonian([2], 3)
    R.add_term(np.array([[1.0, 2.0, 0.0], [2.0, 0.0, 1.0], [0.0, 1.0, -1.0]]))
    R.build("sparse")
    assert not R.Ham.is_complex and np.array_equal(R.Ham.to_scipy().toarray(),
                                                   np.array([[1.0, 2.0, 0.0], [2.0, 0.0, 1.0], [0.0, 1.0, -1.0]]))
    with pytest.raises(ValueError):
        R.build("csr")
    with pytest.raises(TypeError):
        R.add_term([1, 2, 3])
