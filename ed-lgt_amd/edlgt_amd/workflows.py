"""Parameter-dictionary drivers with the reference's layout
(edlgt/workflows/qed.py, edlgt/workflows/su2.py): the same `par` keys, the
same result dictionaries, on the device engine.  Scripts and validation
cases written for the reference's run_QED_spectrum / run_QED_dynamics /
run_SU2_spectrum / run_SU2_dynamics run unchanged with these imports."""
import logging
from time import perf_counter

import numpy as np

from .models import QED_Model, SU2_Model, diagonalize_density_matrix

logger = logging.getLogger(__name__)

__all__ = ["run_QED_spectrum", "run_QED_dynamics", "run_SU2_spectrum", "run_SU2_dynamics",
           "check_observables", "su2_get_momentum_params"]


def _get(d, path, default=None):
    """Nested lookup with a default (workflows/qed.py:12-18)."""
    cur = d
    for k in path:
        if not isinstance(cur, dict) or k not in cur:
            return default
        cur = cur[k]
    return cur


def _get_dtype_mode(par):
    """hamiltonian.dtype_mode, or the legacy hamiltonian.is_complex flag
    (workflows/qed.py:21-27)."""
    mode = _get(par, ["hamiltonian", "dtype_mode"], "auto")
    if mode == "auto":
        legacy = _get(par, ["hamiltonian", "is_complex"], None)
        if legacy is not None:
            mode = bool(legacy)
    return mode


def _set_momentum(model, par):
    if _get(par, ["momentum", "get_momentum_basis"], False):
        model.set_momentum_sector(_get(par, ["momentum", "unit_cell_size"], None),
                                  _get(par, ["momentum", "momentum_k_vals"], None),
                                  _get(par, ["momentum", "TC_symmetry"], False))


def _time_line(par):
    start, stop, dt = par["dynamics"]["start"], par["dynamics"]["stop"], par["dynamics"]["delta_n"]
    return np.arange(start, stop + dt, dt)


def _entropy_and_configs(model, st, res, flags, ii):
    """Entropy / RDM / dominant configurations of one state, sector basis only
    (workflows/qed.py:123-139)."""
    if model.momentum_basis is not None:
        return
    if flags["get_rdm"]:
        rdm = st.reduced_density_matrix(flags["partition_indices"], model._partition_cache)
        ev, _ = diagonalize_density_matrix(rdm)
        res["eigvals"] = ev[np.argsort(ev)[::-1]]
    if flags["get_entropy"]:
        res["entropy"][ii] = st.entanglement_entropy(flags["partition_indices"], model._partition_cache)
    if flags["get_state_configs"]:
        st.get_state_configurations(1e-3, model.sector_configs)


def _partition_flags(model, par, res, n_points):
    partition = _get(par, ["observables", "entropy_partition"], [])
    get_entropy = _get(par, ["observables", "get_entropy"], False)
    get_rdm = _get(par, ["observables", "get_RDM"], False)
    if get_entropy or get_rdm:
        model._get_partition(partition)
        res["entropy"] = np.zeros(n_points, dtype=float)
    return dict(measure_obs=_get(par, ["observables", "measure_obs"], False), get_entropy=get_entropy,
                get_rdm=get_rdm, get_overlap=_get(par, ["observables", "get_overlap"], False),
                get_state_configs=_get(par, ["observables", "get_state_configs"], False),
                partition_indices=partition)


def _plaquette_names(dim, theta):
    if dim == 2:
        return [["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"]]
    if dim == 3:
        out = [["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"], ["C_px,pz", "C_pz,mx", "C_mz,px", "C_mx,mz"],
               ["C_py,pz", "C_pz,my", "C_mz,py", "C_my,mz"]]
        if abs(theta) > 1e-10:
            out += [["EzC_px,py", "C_py,mx", "C_my,px", "C_mx,my"], ["EyC_px,pz", "C_pz,mx", "C_mz,px", "C_mx,mz"],
                    ["ExC_py,pz", "C_pz,my", "C_mz,py", "C_my,mz"]]
        return out
    return []


def _overlap_state(model, par, key_path):
    name = _get(par, key_path, None)
    return model.get_qmb_state_from_configs([model.overlap_QMB_state(name)])


# ------------------------------------------------------------------ QED
def QED_build_model_and_hamiltonian(par):
    """workflows/qed.py:30-45."""
    model = QED_Model(**par["model"])
    _set_momentum(model, par)
    model.default_params()
    m = par.get("m", None) if not model.pure_theory else None
    model.build_Hamiltonian(par["g"], m, par.get("theta", 0), dtype_mode=_get_dtype_mode(par))
    return model


def QED_prepare_observables(model, par, n_points):
    """workflows/qed.py:48-103."""
    theta = par.get("theta", 0.0) if model.pure_theory else 0
    local_obs = [f"E2_p{d}" for d in model.directions]
    if not model.pure_theory:
        local_obs += ["N", "N_zero"]
    plaquette_obs = _plaquette_names(model.dim, theta)
    model.get_observables(local_obs, [], plaquette_obs, twobody_axes=[])
    res = {"E2": np.zeros(n_points), "N": np.zeros(n_points)}
    for names in plaquette_obs:
        res["_".join(names)] = np.zeros(n_points)
    flags = _partition_flags(model, par, res, n_points)
    flags.update(local_obs=local_obs, plaquette_obs=plaquette_obs)
    return res, flags


def QED_measure_on_states(model, par, res, flags, n_points, state_getter, overlap_state=None, dynamics=False):
    """workflows/qed.py:106-162."""
    for ii in range(n_points):
        st = state_getter(ii)
        if not dynamics:
            model.H.print_energy(ii)
        _entropy_and_configs(model, st, res, flags, ii)
        if flags["measure_obs"]:
            model.measure_observables(ii, dynamics=dynamics)
            res["E2"][ii] = model.link_avg(obs_name="E2")
            if not model.pure_theory:
                res["N"][ii] += 0.5 * model.stag_avg("N", "even")
                res["N"][ii] += 0.5 * model.stag_avg("N_zero", "odd")
            for names in flags["plaquette_obs"]:
                obs = "_".join(names)
                res[obs][ii] = model.res[obs]
        if overlap_state is not None and flags["get_overlap"]:
            res["overlap"][ii] = model.measure_fidelity(overlap_state, ii, print_value=True, dynamics=dynamics)


def run_QED_spectrum(par):
    """workflows/qed.py:165-201."""
    t0 = perf_counter()
    model = QED_build_model_and_hamiltonian(par)
    n_eigs = _get(par, ["hamiltonian", "n_eigs"], "full")
    model.diagonalize_Hamiltonian(n_eigs, model.ham_format, print_results=True)
    n_points = len(model.H.Nenergies)  # n_eigs, or every state for 'full'
    res, flags = QED_prepare_observables(model, par, n_points)
    res["energy"] = model.H.Nenergies
    if _get(par, ["hamiltonian", "save_psi"], False):
        for ii in range(n_points):
            res[f"psi{ii}"] = model.H.Npsi[ii].psi
    overlap_state = None
    if flags["get_overlap"]:
        overlap_state = _overlap_state(model, par, ["hamiltonian", "state"])
        res["overlap"] = np.zeros(n_points)
    QED_measure_on_states(model, par, res, flags, n_points, lambda ii: model.H.Npsi[ii], overlap_state, False)
    res["total_time"] = perf_counter() - t0
    logger.info(f"TIME SIMS {res['total_time']:.5f}")
    return res


def run_QED_dynamics(par):
    """workflows/qed.py:204-242."""
    t0 = perf_counter()
    model = QED_build_model_and_hamiltonian(par)
    time_line = _time_line(par)
    n_points = len(time_line)
    res, flags = QED_prepare_observables(model, par, n_points)
    res["time_steps"] = time_line
    in_state = _overlap_state(model, par, ["dynamics", "state"])
    overlap_state = in_state if flags["get_overlap"] else None
    if overlap_state is not None:
        res["overlap"] = np.zeros(n_points)
    if par["dynamics"]["time_evolution"]:
        model.time_evolution_Hamiltonian(in_state, time_line)
    QED_measure_on_states(model, par, res, flags, n_points, lambda ii: model.H.psi_time[ii], overlap_state, True)
    res["total_time"] = perf_counter() - t0
    logger.info(f"TIME SIMS {res['total_time']:.5f}")
    return res


def check_observables(res, ref, atol=1e-10, tag=""):
    """workflows/qed.py:245-254: res[key][0] against every non-None ref value."""
    for key, val in ref.items():
        if val is None:
            continue
        if key not in res:
            raise KeyError(f"{tag}: missing observable '{key}'")
        got = float(res[key][0])
        if not abs(got - float(val)) < atol:
            logger.info(f"{tag}: {key} expected {val} got {got}")
            raise ValueError(f"{tag}: FAIL on observable '{key}'")


# ------------------------------------------------------------------ SU(2)
def su2_build_model_and_hamiltonian(par):
    """workflows/su2.py:92-107."""
    model = SU2_Model(**par["model"])
    _set_momentum(model, par)
    model.default_params()
    m = par.get("m", None) if not model.pure_theory else None
    theta = par.get("theta", 0.0) if model.pure_theory else 0.0
    model.build_Hamiltonian(par["g"], m, theta=theta, dtype_mode=_get_dtype_mode(par))
    return model


def su2_prepare_observables(model, par, n_points):
    """workflows/su2.py:110-170."""
    theta = par.get("theta", 0.0) if model.pure_theory else 0
    local_obs = [f"T2_p{d}" for d in model.directions]
    if not model.pure_theory:
        local_obs += [f"N_{label}" for label in ["tot", "single", "pair", "zero"]]
    plaquette_obs = _plaquette_names(model.dim, theta) if model.dim in (2, 3) else []
    model.get_observables(local_obs, [], plaquette_obs, twobody_axes=[])
    res = {"E2": np.zeros(n_points)}
    for obs in local_obs:
        res[obs] = np.zeros(n_points)
    for names in plaquette_obs:
        res["_".join(names)] = np.zeros(n_points)
    flags = _partition_flags(model, par, res, n_points)
    flags.update(local_obs=local_obs, twobody_obs=[], plaquette_obs=plaquette_obs)
    return res, flags


def su2_measure_on_states(model, par, res, flags, n_points, state_getter, overlap_state=None, dynamics=False):
    """workflows/su2.py:173-227."""
    for ii in range(n_points):
        st = state_getter(ii)
        if not dynamics:
            model.H.print_energy(ii)
        _entropy_and_configs(model, st, res, flags, ii)
        if flags["measure_obs"]:
            model.measure_observables(ii, dynamics=dynamics)
            res["E2"][ii] = model.link_avg(obs_name="T2")
            if not model.pure_theory:
                res["N_single"][ii] = model.stag_avg("N_single")
                res["N_pair"][ii] += 0.5 * model.stag_avg("N_pair", "even")
                res["N_pair"][ii] += 0.5 * model.stag_avg("N_zero", "odd")
                res["N_zero"][ii] += 0.5 * model.stag_avg("N_zero", "even")
                res["N_zero"][ii] += 0.5 * model.stag_avg("N_pair", "odd")
                res["N_tot"][ii] = res["N_single"][ii] + 2.0 * res["N_pair"][ii]
            for names in flags["plaquette_obs"]:
                obs = "_".join(names)
                res[obs][ii] = model.res[obs]
        if overlap_state is not None and flags["get_overlap"]:
            res["overlap"][ii] = model.measure_fidelity(overlap_state, ii, print_value=True, dynamics=dynamics)
    return res


def run_SU2_spectrum(par):
    """workflows/su2.py:230-270."""
    t0 = perf_counter()
    model = su2_build_model_and_hamiltonian(par)
    n_eigs = _get(par, ["hamiltonian", "n_eigs"], "full")
    model.diagonalize_Hamiltonian(n_eigs, model.ham_format, print_results=True)
    n_points = len(model.H.Nenergies)  # n_eigs, or every state for 'full'
    res, flags = su2_prepare_observables(model, par, n_points)
    res["energy"] = model.H.Nenergies
    if _get(par, ["hamiltonian", "save_psi"], False):
        for ii in range(n_points):
            res[f"psi{ii}"] = model.H.Npsi[ii].psi
    overlap_state = None
    if flags["get_overlap"]:
        overlap_state = _overlap_state(model, par, ["hamiltonian", "state"])
        res["overlap"] = np.zeros(n_points)
    res = su2_measure_on_states(model, par, res, flags, n_points, lambda ii: model.H.Npsi[ii], overlap_state, False)
    res["total_time"] = perf_counter() - t0
    logger.info(f"TIME SIMS {res['total_time']:.5f}")
    return res


def run_SU2_dynamics(par):
    """workflows/su2.py:273-313."""
    t0 = perf_counter()
    model = su2_build_model_and_hamiltonian(par)
    time_line = _time_line(par)
    n_points = len(time_line)
    res, flags = su2_prepare_observables(model, par, n_points)
    res["time_steps"] = time_line
    in_state = _overlap_state(model, par, ["dynamics", "state"])
    overlap_state = in_state if flags["get_overlap"] else None
    if overlap_state is not None:
        res["overlap"] = np.zeros(n_points)
    if par["dynamics"]["time_evolution"]:
        model.time_evolution_Hamiltonian(in_state, time_line)
    res = su2_measure_on_states(model, par, res, flags, n_points, lambda ii: model.H.psi_time[ii], overlap_state, True)
    res["total_time"] = perf_counter() - t0
    logger.info(f"TIME SIMS {res['total_time']:.5f}")
    return res


def su2_get_momentum_params(TC_symmetry, n_sites):
    """workflows/su2.py:316-337."""
    if TC_symmetry:
        n_momenta, uc = n_sites, [1]
        k_physical = 2 * np.pi * np.arange(n_momenta) / n_sites
    else:
        n_momenta, uc = n_sites // 2, [2]
        k_physical = 4 * np.pi * np.arange(n_momenta) / n_sites
    return {"n_momenta": n_momenta, "k_indices": np.arange(0, n_momenta, 1), "k_physical": k_physical,
            "k_unit_cell_size": uc, "TC_symmetry": TC_symmetry}
