"""GPU: observables (diagonal locals via site marginals, plaquettes via the
device expectation pass over all plaquettes) against the reference's KAT
observables, and
Krylov exp(-iHt) against the oracle's expm_multiply / exact dense propagation."""
import numpy as np
import pytest

from oracle.model import QEDOracle, expm_path

pytestmark = pytest.mark.gpu


def _qed_observables(mod):
    """workflows/qed.py:48-162 aggregation: E2 = link_avg('E2'), N = 0.5 even(N) + 0.5 odd(N_zero)."""
    mod.get_observables(local_obs=[f"E2_p{d}" for d in mod.directions] + ["N", "N_zero"],
                        plaquette_obs=[["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"]] if mod.dim == 2 else [])
    mod.measure_observables(0)
    out = {"E2": mod.link_avg("E2"), "N": 0.5 * mod.stag_avg("N", "even") + 0.5 * mod.stag_avg("N_zero", "odd")}
    if mod.dim == 2:
        out["plaq"] = mod.res["C_px,py_C_py,mx_C_my,px_C_mx,my"]
    return out


@pytest.mark.parametrize("obc,ref", [
    ([True, True], {"E2": 0.11506337997036731, "N": 0.06869112048269635, "plaq": 0.35977232322995706}),
    ([False, False], {"E2": 0.1906093125650072, "N": 0.11057207253261524, "plaq": 0.36750298319023755}),
])
def test_qed_2d_kat_observables(gpu, obc, ref):
    """validation/qed/test_qed_2d_matter.py:33-38, 67-72 (atol 1e-10)."""
    from edlgt_amd import QED_Model
    mod = QED_Model(spin=1, pure_theory=False, lvals=[4, 2], has_obc=obc)
    mod.build_Hamiltonian(1.0, 1.0)
    mod.diagonalize_Hamiltonian(1, "sparse")
    got = _qed_observables(mod)
    for k, v in ref.items():
        assert abs(got[k] - v) < 1e-10, (k, got[k], v)


def test_qed_1d_kat_observables(gpu):
    """validation/qed/test_qed_1d_matter.py:33-37 pins E2 and N for L=4 OBC (atol 1e-10)."""
    from edlgt_amd import QED_Model
    mod = QED_Model(spin=1, pure_theory=False, lvals=[4], has_obc=[True])
    mod.build_Hamiltonian(1.0, 1.0)
    mod.diagonalize_Hamiltonian(1, "sparse")
    got = _qed_observables(mod)
    assert abs(got["E2"] - 0.03161445624423755) < 1e-10
    assert abs(got["N"] - 0.0474013240232069) < 1e-10


def test_krylov_matches_expm_multiply(gpu):
    import torch
    from edlgt_amd import QED_Model
    lv, obc = [10], [True]
    mod = QED_Model(spin=1, pure_theory=False, lvals=lv, has_obc=obc)
    mod.build_Hamiltonian(1.0, 1.0)
    orc = QEDOracle(1, False, lv, obc)
    Ho = orc.hamiltonian(1.0, 1.0)
    rng = np.random.default_rng(3)
    psi0 = rng.standard_normal(orc.N) + 1j * rng.standard_normal(orc.N)
    psi0 /= np.linalg.norm(psi0)
    tl = np.linspace(0.0, 2.0, 41)
    mod.time_evolution_Hamiltonian(psi0, tl, m=30)
    ref = expm_path(Ho, psi0, tl[0], tl[-1], tl.size)
    got = np.stack([mod.H.psi_time[k].psi for k in range(tl.size)])
    assert np.max(np.abs(got - ref)) < 1e-10
    # per-step diagonal observable from marginals vs dense
    nvals = np.real(np.einsum("sii->si", orc.ops["N"]))
    dens = np.abs(ref) ** 2
    expect = dens @ np.stack([nvals[s, orc.configs[:, s]] for s in range(orc.n_sites)], axis=1)
    got_obs = np.einsum("tsa,sa->ts", mod.H.marginals, nvals[:, : mod.H.marginals.shape[2]])
    assert np.max(np.abs(got_obs - expect)) < 1e-9
    assert np.max(mod.H.krylov_err) < 1e-8


def test_krylov_mixed_precision(gpu):
    from edlgt_amd import QED_Model, PREC_MIXED
    lv, obc = [4, 2], [True, True]
    mod = QED_Model(spin=1, pure_theory=False, lvals=lv, has_obc=obc)
    mod.build_Hamiltonian(1.0, 1.0)
    orc = QEDOracle(1, False, lv, obc)
    Ho = orc.hamiltonian(1.0, 1.0)
    psi0 = np.zeros(orc.N, dtype=complex)
    psi0[0] = 1.0
    tl = np.linspace(0.0, 1.0, 21)
    mod.time_evolution_Hamiltonian(psi0, tl, m=30, precision=PREC_MIXED)
    ref = expm_path(Ho, psi0, 0.0, 1.0, 21)
    got = np.stack([mod.H.psi_time[k].psi for k in range(21)])
    assert np.max(np.abs(got - ref)) < 1e-5
