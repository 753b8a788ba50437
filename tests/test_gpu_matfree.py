"""GPU: the matrix-free operator (edlgt_hamiltonian_build_matfree) against the
stored CSR of the same program: products (1e-13 relative), lowest
eigenvalues (1e-10), Krylov steps (1e-10), the reference's known answer
through the model switch, and the entry points that need stored entries
refusing it."""
import numpy as np
import pytest

from oracle.model import QEDOracle

pytestmark = pytest.mark.gpu




@pytest.mark.parametrize("kind", ["qed2x2", "su2_l6", "z3_3x3"])
def test_products_and_spectra_match_stored(gpu, kind):
    import torch
    from edlgt_amd import QED_Model, SU2_Model, Zn_Model

    def make():
        if kind == "qed2x2":
            return QED_Model(spin=1, pure_theory=False, lvals=[2, 2], has_obc=[False, False]), (0.7, 0.3)
        if kind == "su2_l6":
            return SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False]), (1.0, 5.0)
        return Zn_Model(3, lvals=[3, 3], has_obc=[False, False], flux_sector=(0, 0)), (1.0, 0.7)

    ms, args = make()
    ms.matrix_free = False
    Hs = ms.build_Hamiltonian(*args)
    mf, _ = make()
    mf.matrix_free = True
    Hf = mf.build_Hamiltonian(*args)
    assert Hf.matrix_free and not Hs.matrix_free and Hf.nnz == 0
    rng = np.random.default_rng(2)
    dt = torch.complex128 if Hs.is_complex else torch.float64
    x = rng.standard_normal(Hs.n) + (1j * rng.standard_normal(Hs.n) if Hs.is_complex else 0)
    x = torch.as_tensor(x, dtype=dt, device=gpu)
    ys, yf = Hs.matvec(x).cpu().numpy(), Hf.matvec(x).cpu().numpy()
    assert np.max(np.abs(ys - yf)) <= 1e-13 * np.max(np.abs(ys))
    X = torch.stack([x, 2 * x, -x]).contiguous()
    Yf = Hf.matvec(X).cpu().numpy()
    assert np.max(np.abs(Yf[1] - 2 * ys)) <= 1e-13 * np.max(np.abs(ys))
    es, _, _, _ = Hs.lanczos(3, tol=1e-12, want_vectors=False)
    ef, _, _, _ = Hf.lanczos(3, tol=1e-12, want_vectors=False)
    assert np.max(np.abs(es - ef)) < 1e-10


def test_model_switch_kat_and_krylov(gpu):
    """QED 1D open chain of 10 sites (197 states) through the model with
    matrix_free=True: the lowest eigenvalue against the oracle, and Krylov
    steps (states and per-step marginals) equal to the stored operator's."""
    from edlgt_amd import QED_Model
    a = QED_Model(spin=1, pure_theory=False, lvals=[10], has_obc=[True])
    a.matrix_free = True
    a.build_Hamiltonian(1.0, 1.0)
    assert a.H.Ham.matrix_free
    a.diagonalize_Hamiltonian(1, "sparse")
    orc = QEDOracle(1, False, [10], [True])
    from oracle.model import lowest_eigs
    assert abs(a.H.energies[0] - lowest_eigs(orc.hamiltonian(1.0, 1.0), 1)[0]) < 1e-10
    b = QED_Model(spin=1, pure_theory=False, lvals=[10], has_obc=[True])
    b.matrix_free = False
    b.build_Hamiltonian(1.0, 1.0)
    psi0 = a.get_qmb_state_from_configs([a.sector_configs[0]])
    t = np.arange(11) * 0.1
    a.time_evolution_Hamiltonian(psi0, t, store_states=False)
    b.time_evolution_Hamiltonian(psi0, t, store_states=False)
    assert float((a.H.psi_final - b.H.psi_final).abs().max()) < 1e-10
    assert np.max(np.abs(a.H.marginals - b.H.marginals)) < 1e-10


def test_stored_only_entry_points_refuse(gpu):
    from edlgt_amd import Server, SU2_Model
    m = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    m.matrix_free = True
    H = m.build_Hamiltonian(1.0, 5.0)
    for call in (H.to_host, H.compress, H.autotune):
        with pytest.raises(ValueError):
            call()
    with pytest.raises(ValueError):
        H.lanczos(1, precision=1)  # FP32 products need stored values
    with pytest.raises(ValueError):
        Server(H, n_eigs=1)


@pytest.mark.parametrize("kind", ["qed2x2", "su2_l6", "z3_3x3"])
def test_device_count_equals_stored_entries(gpu, kind):
    """edlgt_hamiltonian_count (transition tables only, no enumeration) gives
    the stored CSR's entry count exactly for these models."""
    from edlgt_amd import QED_Model, SU2_Model, Zn_Model
    from edlgt_amd.device import stored_entries
    if kind == "qed2x2":
        m, args = QED_Model(spin=1, pure_theory=False, lvals=[2, 2], has_obc=[False, False]), (0.7, 0.3)
    elif kind == "su2_l6":
        m, args = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False]), (1.0, 5.0)
    else:
        m, args = Zn_Model(3, lvals=[3, 3], has_obc=[False, False], flux_sector=(0, 0)), (1.0, 0.7)
    m.matrix_free = False
    H = m.build_Hamiltonian(*args)
    assert stored_entries(m.basis, np.stack(m._tables), m._placements) == H.nnz


@pytest.mark.parametrize("mode", ["stored", "coded", "free"])
def test_automatic_storage_choice(gpu, monkeypatch, mode):
    """matrix_free = "auto" under a pretended amount of free device memory:
    plenty -> stored CSR; room for the full build but not for a solver beside
    it -> dictionary-coded CSR with its values released; too little -> the
    matrix-free operator.  Every form gives the stored spectrum (QED 2x2 PBC,
    547 states)."""
    import torch
    from edlgt_amd import QED_Model
    ref = QED_Model(spin=1, pure_theory=False, lvals=[2, 2], has_obc=[False, False])
    ref.matrix_free = False
    Hr = ref.build_Hamiltonian(0.7, 0.3)
    er, _, _, _ = Hr.lanczos(3, tol=1e-12, want_vectors=False)
    full = Hr.nnz * (4 + 16) + 8 * (Hr.n + 1)  # complex values
    free = {"stored": 10 * full, "coded": int(1.5 * full), "free": full // 2}[mode]
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda *a: (free, 300 * 10**9))
    m = QED_Model(spin=1, pure_theory=False, lvals=[2, 2], has_obc=[False, False])
    H = m.build_Hamiltonian(0.7, 0.3)
    got = "free" if H.matrix_free else ("coded" if getattr(H, "n_dict", 0) > 0 else "stored")
    assert got == mode
    e, _, _, _ = H.lanczos(3, tol=1e-12, want_vectors=False)
    assert np.max(np.abs(e - er)) < 1e-10
