"""GPU: 1D QED with the gauge field integrated out (QED_Model(spin=
"integrated"), QED_model.py:64-82, 396-652).

The reference's own check (validation/qed/test_qed_1d_integrated.py): the
model's Hamiltonian equals the direct construction (g/2) sum_n [sum_{k<=n}
(q_k + (Sz_k + (-1)^k)/2)]^2 + mass + hopping within 1e-10, and the link
<E^2> reconstructed from measured <N_k> and <N_k N_l> equals the direct
<E_n^2> within 1e-10.  Here against oracle/integrated.py (that direct
construction, pinned by free fermions and Gauss's law) at L = 4 (the KAT's
couplings and charges), 6, 8 and 12.
"""
import numpy as np
import pytest

from oracle import integrated as iq

pytestmark = pytest.mark.gpu

CASES = [
    (4, 1.4, 0.7, [1.0, -1.0, 0.0, 0.0]),  # validation/qed/test_qed_1d_integrated.py
    (6, 1.0, 0.5, None),
    (8, 0.6, 1.2, [0.0, 0.0, 1.0, 0.0, 0.0, -1.0, 0.0, 0.0]),
    (12, 2.0, 0.3, None),
]


@pytest.mark.parametrize("L,g,m,q", CASES)
def test_integrated_hamiltonian_energy_e2(gpu, L, g, m, q):
    from edlgt_amd import QED_Model
    print(f"integrated QED L={L} g={g} m={m} q={q}", flush=True)
    mod = QED_Model(spin="integrated", pure_theory=False, lvals=[L], has_obc=[True], bg_list=q)
    qq = np.zeros(L) if q is None else np.asarray(q, float)
    assert np.array_equal(mod.sector_configs, iq.sector_configs(L))
    assert list(mod.loc_dims) == [2] * L
    H = mod.build_Hamiltonian(g, m)
    Ho = iq.hamiltonian(L, g, m, qq).toarray()
    # the reference's own check reads the dense matrix: np.asarray(model.H.Ham)
    Hg = np.asarray(mod.H.Ham, dtype=np.complex128)
    assert np.max(np.abs(Hg - Ho)) <= 1e-12 * max(1.0, np.max(np.abs(Ho)))
    w, v = np.linalg.eigh(Ho)
    mod.diagonalize_Hamiltonian(1, "sparse")
    assert abs(mod.H.Nenergies[0] * L - w[0]) < 1e-10
    mod.get_observables(local_obs=["N"])
    mod.measure_observables(0)
    e2 = mod.reconstruct_integrated_E2_from_N(state_index=0)
    ref = iq.link_e2(L, qq, mod.H.Npsi[0].psi)
    assert e2.shape == (L - 1,) and np.max(np.abs(e2 - ref)) < 1e-10
    assert mod.res["E2_avg"] == pytest.approx(float(np.mean(ref)), abs=1e-10)
    # N_N measured through the observables API gives the same reconstruction
    mod.res.pop("N_N")
    mod.get_observables(local_obs=["N"], twobody_obs=[["N", "N"]])
    mod.measure_observables(0)
    assert np.max(np.abs(mod.reconstruct_integrated_E2_from_N(print_values=False) - ref)) < 1e-10


def test_integrated_labels_and_arguments(gpu):
    from edlgt_amd import QED_Model
    L = 8
    mod = QED_Model(spin="integrated", pure_theory=False, lvals=[L], has_obc=[True])
    rows = {tuple(c): r for r, c in enumerate(iq.sector_configs(L).tolist())}
    vac = mod.overlap_QMB_state("V")
    assert list(vac) == [1, 0, 1, 0, 1, 0, 1, 0] and tuple(vac) in rows
    assert list(mod.overlap_QMB_state("PV")) == list(vac)
    mes = mod.overlap_QMB_state("meson_center")
    assert list(mes) == [1, 0, 1, 0, 0, 1, 1, 0] and tuple(mes) in rows
    psi = mod.get_qmb_state_from_configs([mes])
    assert int(psi.abs().argmax().item()) == rows[tuple(mes)]
    with pytest.raises(ValueError):
        QED_Model(spin="half", pure_theory=False, lvals=[4], has_obc=[True])
    with pytest.raises(ValueError):
        QED_Model(spin="integrated", pure_theory=False, lvals=[2, 2], has_obc=[True, True])
    with pytest.raises(ValueError):
        QED_Model(spin="integrated", pure_theory=True, lvals=[4], has_obc=[True])
    with pytest.raises(ValueError):
        mod.build_Hamiltonian(1.0)  # m is required


def test_integrated_coupling_sweep_on_one_pattern(gpu):
    """H(g) = (g/2) H_E + H_matter(m) assembled once (build_multi_term) equals
    the oracle's direct matrix at every g."""
    from edlgt_amd import QED_Model
    L, m = 8, 0.9
    q = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, -1.0, 0.0]
    mod = QED_Model(spin="integrated", pure_theory=False, lvals=[L], has_obc=[True], bg_list=q)
    H3 = mod.build_multi_term(m)
    assert H3.n_terms == 3 and set(np.unique(H3.terms())) <= {0, 2}
    for g in (0.25, 1.0, 3.0):
        Ho = iq.hamiltonian(L, g, m, np.asarray(q)).toarray()
        Hm = H3.to_scipy(coef=mod.multi_term_coefficients(g)).toarray()
        assert np.max(np.abs(Hm - Ho)) <= 1e-12 * max(1.0, np.max(np.abs(Ho)))
