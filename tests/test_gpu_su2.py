"""GPU parity for SU(2) (spin-1/2 rishons): sector bit-exact, CSR pattern
bit-exact and values within 1e-14 against the oracle; the reference's known
answers (L=6 PBC, N_tot=6, g=1, m=5: energy, E2, N_*, entropy at 1e-10) and
golden dynamics (F, E2, S at 1e-6, the reference's tolerance), plus the Krylov
path against the oracle's expm_multiply at 1e-8 and the mixed-precision
(complex64 SpMV, FP64 recurrence) path against the FP64 one."""
from pathlib import Path

import numpy as np
import pytest

from oracle.model import SU2Oracle, expm_path, lowest_eigs

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).parent / "golden" / "su2_dynamics_L6_pbc.npz"


def _model(lv, obc, pure=False, n=None):
    from edlgt_amd import SU2_Model
    return SU2_Model(spin=0.5, pure_theory=pure, sectors=None if pure else [n], lvals=lv, has_obc=obc)


def _obs(mod, marg_or_index, dynamics=False):
    mod.measure_observables(marg_or_index, dynamics=dynamics)
    r = {"E2": mod.link_avg("T2"), "N_single": mod.stag_avg("N_single")}
    r["N_pair"] = 0.5 * mod.stag_avg("N_pair", "even") + 0.5 * mod.stag_avg("N_zero", "odd")
    r["N_zero"] = 0.5 * mod.stag_avg("N_zero", "even") + 0.5 * mod.stag_avg("N_pair", "odd")
    r["N_tot"] = r["N_single"] + 2 * r["N_pair"]
    return r


@pytest.mark.parametrize("lv,obc,pure,n", [([6], [False], False, 6), ([10], [True], False, 10),
                                           ([2, 2], [True, True], False, 4), ([2, 2], [False, False], True, None)])
def test_su2_sector_csr_energy(gpu, lv, obc, pure, n):
    print(f"SU2 lvals={lv} obc={obc} pure={pure}", flush=True)
    orc = SU2Oracle(pure, lv, obc, sectors=None if pure else [n])
    mod = _model(lv, obc, pure, n)
    assert mod.sector_dim == orc.N
    assert np.array_equal(mod.sector_configs, orc.configs)
    Hg = mod.build_Hamiltonian(1.0, None if pure else 5.0)
    Ho = orc.hamiltonian(1.0, None if pure else 5.0)
    rp, col, val = Hg.to_host()
    assert np.array_equal(rp, Ho.indptr) and np.array_equal(col, Ho.indices)
    assert np.max(np.abs(val - Ho.data)) <= 1e-14 * np.max(np.abs(Ho.data))
    k = min(4, orc.N - 1)
    mod.diagonalize_Hamiltonian(k, "sparse", tol=1e-12)
    assert np.max(np.abs(mod.H.Nenergies * np.prod(lv) - lowest_eigs(Ho, k))) < 1e-10


def test_su2_kat_observables_entropy(gpu):
    mod = _model([6], [False], False, 6)
    mod.build_Hamiltonian(1.0, 5.0)
    mod.diagonalize_Hamiltonian(1, "sparse", tol=1e-12)
    assert abs(mod.H.Nenergies[0] - (-6.173752796132477)) < 1e-10
    mod.get_observables([f"T2_p{d}" for d in mod.directions] + [f"N_{l}" for l in ("tot", "single", "pair", "zero")])
    r = _obs(mod, 0)
    ref = {"E2": 0.0553604161539725, "N_tot": 0.15535989251437904, "N_single": 0.14157936434695081,
           "N_pair": 0.006890264083714107, "N_zero": 0.8515303715693359}
    for k, v in ref.items():
        assert abs(r[k] - v) < 1e-10, k
    mod._get_partition([0, 1, 2])
    S = mod.H.Npsi[0].entanglement_entropy([0, 1, 2], mod._partition_cache)
    assert abs(S - 0.802957045143) < 1e-10


def test_su2_gold_dynamics(gpu):
    import torch
    from edlgt_amd import PREC_MIXED
    gold = np.load(GOLD, allow_pickle=False)
    t = gold["time_steps"]
    mod = _model([6], [False], False, 6)
    mod.build_Hamiltonian(1.0, 5.0)
    mod.get_observables(["T2_px"] + [f"N_{l}" for l in ("tot", "single", "pair", "zero")])
    psi0 = mod.get_qmb_state_from_configs([mod.overlap_QMB_state("V")])
    mod.time_evolution_Hamiltonian(psi0, t)
    mod._get_partition([0, 1, 2])
    F = np.array([abs(torch.vdot(psi0, s.psi_device).item()) ** 2 for s in mod.H.psi_time])
    E2 = np.array([_obs(mod, i, dynamics=True)["E2"] for i in range(len(t))])
    S = np.array([s.entanglement_entropy([0, 1, 2], mod._partition_cache) for s in mod.H.psi_time])
    for got, ref in ((F, gold["F"]), (E2, gold["E2"]), (S, gold["S"])):
        d = got - ref
        assert np.max(np.abs(d)) < 1e-6 and np.sqrt(np.mean(d ** 2)) < 1e-6
    # against the oracle's expm_multiply path at 1e-8
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    p0 = psi0.cpu().numpy()
    path = expm_path(orc.hamiltonian(1.0, 5.0), p0, t[0], t[-1], len(t))
    got = np.stack([s.psi for s in mod.H.psi_time])
    assert np.max(np.abs(got - path)) < 1e-8
    # mixed precision: complex64 SpMV, FP64 recurrence
    mod.time_evolution_Hamiltonian(psi0, t, precision=PREC_MIXED)
    E2m = np.array([_obs(mod, i, dynamics=True)["E2"] for i in range(len(t))])
    assert np.max(np.abs(E2m - E2)) < 1e-6


def test_su2_krylov_without_full_reorth(gpu):
    """The three-term Krylov recurrence (no full reorthogonalisation, used for
    the 20-site quench) against expm_multiply on a 10-site OBC chain."""
    orc = SU2Oracle(False, [10], [True], sectors=[10])
    mod = _model([10], [True], False, 10)
    mod.build_Hamiltonian(1.0, 1.0)
    psi0 = mod.get_qmb_state_from_configs([mod.overlap_QMB_state("V")])
    t = np.arange(21) * 0.05
    mod.time_evolution_Hamiltonian(psi0, t, m=30, reorth=False)
    path = expm_path(orc.hamiltonian(1.0, 1.0), psi0.cpu().numpy(), t[0], t[-1], len(t))
    got = np.stack([s.psi for s in mod.H.psi_time])
    assert np.max(np.abs(got - path)) < 1e-8
    # chunked evolution through the kept Krylov session (each chunk's time line
    # starts at 0: the time line is measured from the state passed in): same
    # bits as one run
    psi_full = mod.H.psi_final.clone()
    marg_full = mod.H.marginals
    mod.time_evolution_Hamiltonian(psi0, t[:11], m=30, reorth=False, store_states=False)
    ks = mod.H._krylov
    mod.time_evolution_Hamiltonian(mod.H.psi_final, t[:11], m=30, reorth=False, store_states=False)
    assert mod.H._krylov is ks
    assert bool((mod.H.psi_final == psi_full).all())
    assert np.array_equal(mod.H.marginals, marg_full[10:])
    mod.H.release_krylov()
    assert mod.H._krylov is None


@pytest.mark.parametrize("obc", [[False], [True]])
def test_su2_correlators_and_energy_closure(gpu, obc):
    """Two-body correlators (all pairs) and n-body hopping expectation values
    from one device pass each, against the oracle's triplet expectation values
    at 1e-12; and the ground-state energy re-assembled from the measured local,
    mass and hopping terms equals the Lanczos eigenvalue (1e-10)."""
    from oracle.model import expval_nbody, twobody_corr
    mod = _model([6], obc, False, 6)
    mod.build_Hamiltonian(1.0, 5.0)
    mod.diagonalize_Hamiltonian(1, "sparse")
    mod.get_observables(local_obs=["E_square", "N-1"], twobody_obs=[["N_tot", "N_single"]],
                        nbody_obs=[["Qpx_dag", "Qmx"]], nbody_dist=[[(1,)]])
    mod.measure_observables(0)
    psi = mod.H.Npsi[0].psi
    orc = SU2Oracle(False, [6], obc, sectors=[6])
    assert np.array_equal(orc.configs, mod.sector_configs)
    args = (orc.configs, orc.keys, orc.loc_dims, psi)
    corr = twobody_corr(orc.ops["N_tot"], orc.ops["N_single"], *args)
    assert np.max(np.abs(mod.res["N_tot_N_single"] - corr)) < 1e-12
    bonds = [(i, i + 1) for i in range(5)] + ([] if obc[0] else [(5, 0)])
    hop = np.array([expval_nbody([orc.ops["Qpx_dag"], orc.ops["Qmx"]], list(b), *args) for b in bonds])
    assert mod.res["Qpx_dag_Qmx"].shape == (len(bonds),)
    assert np.max(np.abs(mod.res["Qpx_dag_Qmx"] - hop.real)) < 1e-12
    # closure: <H> from the measured pieces (SU2_model.py:200-355 term list)
    c = orc.couplings(1.0, 5.0)
    mass = np.array([c["m_even"] if i % 2 == 0 else c["m_odd"] for i in range(6)])
    e = c["E"] * np.sum(mod.res["E_square"]) + np.sum(mass * mod.res["N-1"])
    e += sum(2 * (c["tx_even" if b[0] % 2 == 0 else "tx_odd"] * h).real for b, h in zip(bonds, hop))
    assert abs(e - mod.H.Nenergies[0] * 6) < 1e-10
