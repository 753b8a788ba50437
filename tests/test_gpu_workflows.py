"""GPU: the parameter-dictionary drivers (edlgt_amd.workflows, the layout of
the reference's edlgt/workflows/qed.py and su2.py) on the reference's own
validation cases and known answers: validation/qed/test_qed_1d_matter.py,
test_qed_2d_matter.py, test_qed_2d_pure.py (energy per site and observables
at 1e-10), validation/su2/test_su2_spectrum.py (with entropy) and
test_su2_dynamics.py (F, E2, S against the golden series at 1e-6)."""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).parent / "golden" / "su2_dynamics_L6_pbc.npz"


def _qed_par(lvals, obc, pure, g, m=None):
    par = {"model": {"lvals": lvals, "has_obc": obc, "spin": 1, "pure_theory": pure, "ham_format": "sparse",
                     "bg_list": None},
           "hamiltonian": {"n_eigs": 1, "save_psi": False},
           "momentum": {"get_momentum_basis": False, "unit_cell_size": [1] * len(lvals),
                        "momentum_k_vals": [0] * len(lvals)},
           "observables": {"measure_obs": True, "get_entropy": False, "get_state_configs": False,
                           "get_overlap": False},
           "g": g}
    if m is not None:
        par["m"] = m
    return par


@pytest.mark.parametrize("lvals,obc,pure,m,ref", [
    ([4], [True], False, 1.0, {"energy": -0.5691411466075982, "E2": 0.03161445624423755, "N": 0.0474013240232069}),
    ([4], [False], False, 1.0, {"energy": -0.5903881075118848, "E2": 0.029961500436415017, "N": 0.0598294699532885}),
    ([4, 2], [True, True], False, 1.0, {"energy": -0.6862839217038432, "E2": 0.11506337997036731,
                                         "N": 0.06869112048269635,
                                         "C_px,py_C_py,mx_C_my,px_C_mx,my": 0.35977232322995706}),
    ([4, 2], [False, False], False, 1.0, {"energy": -0.8830087027925408, "E2": 0.1906093125650072,
                                           "N": 0.11057207253261524,
                                           "C_px,py_C_py,mx_C_my,px_C_mx,my": 0.36750298319023755}),
    ([3, 3], [True, True], True, None, {"energy": -0.09443358823578024, "E2": 0.10478557317866576,
                                         "C_px,py_C_py,mx_C_my,px_C_mx,my": 0.36965393329850427}),
    ([3, 3], [False, False], True, None, {"energy": -0.2042469295595554, "E2": 0.14512577296208287,
                                           "C_px,py_C_py,mx_C_my,px_C_mx,my": 0.349372702521638}),
])
def test_run_qed_spectrum_kats(gpu, lvals, obc, pure, m, ref):
    from edlgt_amd.workflows import check_observables, run_QED_spectrum
    res = run_QED_spectrum(_qed_par(lvals, obc, pure, 1.0, m))
    check_observables(res, ref, atol=1e-10, tag=f"QED {lvals} {obc}")
    with pytest.raises(ValueError):
        check_observables(res, {"energy": ref["energy"] + 1e-6}, atol=1e-10)


def _su2_par():
    return {"model": {"lvals": [6], "sectors": [6], "has_obc": [False], "spin": 0.5, "pure_theory": False,
                      "ham_format": "sparse"},
            "momentum": {"get_momentum_basis": False, "unit_cell_size": [2], "TC_symmetry": False},
            "g": 1.0, "m": 5.0}


def test_run_su2_spectrum_kat(gpu):
    from edlgt_amd.workflows import run_SU2_spectrum
    par = _su2_par()
    par["hamiltonian"] = {"n_eigs": 1, "save_psi": True}
    par["observables"] = {"measure_obs": True, "get_entropy": True, "entropy_partition": [0, 1, 2],
                          "get_state_configs": True, "get_overlap": False, "get_RDM": True}
    res = run_SU2_spectrum(par)
    ref = {"energy": -6.173752796132477, "entropy": 0.802957045143, "E2": 0.0553604161539725,
           "N_tot": 0.15535989251437904, "N_single": 0.14157936434695081, "N_pair": 0.006890264083714107,
           "N_zero": 0.8515303715693359}
    for k, v in ref.items():
        assert abs(res[k][0] - v) < 1e-10, k
    lam = res["eigvals"]
    assert abs(np.sum(lam) - 1) < 1e-12 and np.all(np.diff(lam) <= 1e-15)
    lam = lam[lam > 1e-10]
    assert abs(-np.sum(lam * np.log2(lam)) - ref["entropy"]) < 1e-10
    assert res["psi0"].shape == (282,)


def test_run_su2_dynamics_gold(gpu):
    from edlgt_amd.workflows import run_SU2_dynamics
    par = _su2_par()
    par["dynamics"] = {"time_evolution": True, "start": 0.0, "stop": 0.5, "delta_n": 0.01, "state": "V"}
    par["observables"] = {"measure_obs": True, "get_entropy": True, "entropy_partition": [0, 1, 2],
                          "get_overlap": True, "get_state_configs": False, "get_RDM": False}
    res = run_SU2_dynamics(par)
    gold = np.load(GOLD, allow_pickle=False)
    assert np.allclose(res["time_steps"], gold["time_steps"], atol=1e-14, rtol=0)
    for got, ref in ((res["overlap"], gold["F"]), (res["E2"], gold["E2"]), (res["entropy"], gold["S"])):
        d = np.asarray(got) - ref
        assert np.max(np.abs(d)) < 1e-6 and np.sqrt(np.mean(d ** 2)) < 1e-6


def test_run_su2_spectrum_in_momentum_sector(gpu):
    """The same case through the momentum-sector path (k = 0, unit cell 2):
    the ground state and its observables are those of the full sector."""
    from edlgt_amd.workflows import run_SU2_spectrum, su2_get_momentum_params
    mp = su2_get_momentum_params(False, 6)
    assert mp["n_momenta"] == 3 and mp["k_unit_cell_size"] == [2]
    par = _su2_par()
    par["momentum"] = {"get_momentum_basis": True, "unit_cell_size": mp["k_unit_cell_size"],
                       "momentum_k_vals": [0], "TC_symmetry": False}
    par["hamiltonian"] = {"n_eigs": 1, "save_psi": False}
    par["observables"] = {"measure_obs": True, "get_entropy": True, "entropy_partition": [0, 1, 2],
                          "get_state_configs": False, "get_overlap": False}
    res = run_SU2_spectrum(par)
    assert abs(res["energy"][0] - (-6.173752796132477)) < 1e-10
    assert abs(res["E2"][0] - 0.0553604161539725) < 1e-10
    assert res["entropy"][0] == 0.0  # not measured in a momentum sector, as in the reference


def test_driver_through_run_sim_result_layout(gpu, tmp_path):
    """A QED KAT case generated from a config template, run through
    run_sim with the spectrum driver and read back with get_sim from
    results/<uid>/res.npz (the reference scripts' simsio flow,
    scripts/script_QED.py:18-40)."""
    import yaml
    from edlgt_amd import simio
    from edlgt_amd.workflows import check_observables, run_QED_spectrum
    common = _qed_par([4], [True], False, 1.0, 1.0)
    del common["g"], common["m"]
    cfg = tmp_path / "configs" / "QED"
    cfg.mkdir(parents=True)
    header = {"===": {"template": "$uid:\n  <<<: common\n  g: $g\n  m: $m\n", "common": common}}
    (cfg / "kat.yaml").write_text(yaml.safe_dump(header, sort_keys=False))
    (uid,) = simio.gen_configs("template", {"g": [1.0], "m": [1.0]}, "QED/kat", tmp_path / "configs")
    with simio.run_sim(uid, tmp_path / "configs", tmp_path / "results") as sim:
        sim.res.update(run_QED_spectrum(sim.par))
    back = simio.get_sim(uid, tmp_path / "configs", tmp_path / "results")
    check_observables(back.res, {"energy": -0.5691411466075982, "E2": 0.03161445624423755,
                                 "N": 0.0474013240232069}, atol=1e-10, tag="QED 1D via run_sim")
    q = simio.SimsQuery("QED/kat", tmp_path / "configs")
    assert q.uids == [uid]
