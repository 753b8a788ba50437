"""GPU parity of the ensemble averages (quantum_model.py:912-1156) on the
fully diagonalised 6-site periodic SU(2) chain with matter (N_tot = 6): the
engine's microcanonical, diagonal and canonical averages against the oracle's
loops over the dense eigendecomposition (oracle/ensembles.py, itself pinned by
closed forms in tests/test_ensembles.py), with the three reductions the
reference offers (site mean, staggered mean, normalisation vector / L), for a
seeded random state and a basis configuration; the diagonal ensemble again
from the real-gauge solve.  Eigenvectors differ from the oracle's by phases
and by rotations inside degenerate levels; every compared quantity is
invariant under both (shells hold whole levels: no eigenvalue lies within
1e-6 of a shell edge, checked)."""
import numpy as np
import pytest

from oracle import ensembles as ens
from oracle import lattice as lat
from oracle.model import SU2Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-9


@pytest.fixture(scope="module")
def setup(gpu):
    from edlgt_amd import SU2_Model
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    H = orc.hamiltonian(1.0, 5.0)
    e, v = np.linalg.eigh(H.toarray())
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    mod.build_Hamiltonian(1.0, 5.0)
    mod.diagonalize_Hamiltonian("full", "sparse", tol=1e-12)
    assert np.max(np.abs(mod.H.energies - e)) < 1e-10
    even = np.array([sum(lat.zig_zag([6], i)) % 2 == 0 for i in range(6)])
    norm = np.array([1.0, -1.0, 1.0, -1.0, 1.0, -1.0])
    diags = {k: np.real(np.einsum("sii->si", orc.ops[k])) for k in ("N_single", "N_pair", "N_tot")}
    post = {"N_single": np.mean, "N_pair": lambda x: float(np.mean(x[even])),
            "N_tot": lambda x: float(np.dot(x, norm) / 6)}
    kw = dict(special_norms={"N_tot": norm}, staggered_avgs={"N_pair": "even"})
    rng = np.random.default_rng(7)
    psi = rng.normal(size=orc.N) + 1j * rng.normal(size=orc.N)
    states = {"random": psi / np.linalg.norm(psi), "config": np.eye(orc.N)[0], "config1": np.eye(orc.N)[1]}
    return orc, H, e, v, mod, diags, post, kw, states


@pytest.mark.parametrize("which", ["random", "config"])
def test_microcanonical(setup, which):
    orc, H, e, v, mod, diags, post, kw, states = setup
    psi = states[which]
    print(f"microcanonical state={which}", flush=True)
    Eq = np.vdot(psi, H @ psi).real
    dE = np.sqrt(np.vdot(H @ psi, H @ psi).real - Eq ** 2)
    assert np.min(np.abs(np.abs(e - Eq) - dE)) > 1e-6
    th_o, want = ens.microcanonical(H, v, e / 6, psi, orc.configs, diags, post, 6)
    th, got = mod.microcanonical_avg(list(diags), psi, **kw)
    for k in want:
        assert abs(got[k] - want[k]) < TOL, k
    assert abs(np.linalg.norm(th) - 1.0) < 1e-10
    # the coherent shell state lies in the shell's span
    inside = np.nonzero(np.abs(e / 6 - Eq / 6) < dE / 6)[0]
    proj = v[:, inside] @ (v[:, inside].conj().T @ th)
    assert np.linalg.norm(proj - th) < 1e-9


@pytest.mark.parametrize("which", ["random", "config"])
def test_diagonal(setup, which):
    orc, H, e, v, mod, diags, post, kw, states = setup
    psi = states[which]
    print(f"diagonal state={which}", flush=True)
    gaps = np.abs(np.diff(e / 6))
    assert np.all((gaps < 1e-12) | (gaps > 1e-6))  # levels split the same way on both sides
    want = ens.diagonal(v, e / 6, psi, orc.configs, diags, post)
    got = mod.diagonal_avg(list(diags), psi, **kw)
    for k in want:
        assert abs(got[k] - want[k]) < TOL, k


@pytest.mark.parametrize("beta", [0.0, 0.3, 2.0, 50.0])
def test_canonical(setup, beta):
    orc, H, e, v, mod, diags, post, kw, states = setup
    print(f"canonical beta={beta}", flush=True)
    for k in diags:
        want = ens.canonical(e, v, orc.configs, diags[k], beta, 6)
        assert abs(mod.canonical_avg(k, beta) - want) < TOL, k


def test_real_gauge_solve_and_partial_spectrum(setup):
    from edlgt_amd import SU2_Model
    orc, H, e, v, mod, diags, post, kw, states = setup
    print("diagonal ensemble from the real-gauge solve", flush=True)
    m2 = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    m2.build_Hamiltonian(1.0, 5.0)
    m2.diagonalize_Hamiltonian("full", "sparse", tol=1e-12, real_gauge=True)
    assert m2.H.real_gauge
    want = ens.diagonal(v, e / 6, states["random"], orc.configs, diags, post)
    got = m2.diagonal_avg(list(diags), states["random"], **kw)
    for k in want:
        assert abs(got[k] - want[k]) < TOL, k
    m2.diagonalize_Hamiltonian(4, "sparse", tol=1e-12)
    with pytest.raises(ValueError):
        m2.diagonal_avg(["N_single"], states["random"])
    # the canonical average needs every level: one device solve of the sector
    for k in diags:
        assert abs(m2.canonical_avg(k, 0.3) - ens.canonical(e, v, orc.configs, diags[k], 0.3, 6)) < TOL, k


def test_spectral_sums_beta_and_r(setup):
    """partition_function / free_energy / thermal_average / F_prime / get_beta
    (qmb_hamiltonian.py:428-564) in both of the reference's branches: over the
    stored per-site energies after the full solve, and, after a partial solve,
    over the eigenvalues of H (the oracle takes the dense exponential's
    traces there); get_r_value; compute_expH against scipy's expm."""
    from scipy.linalg import expm
    from edlgt_amd import SU2_Model
    orc, H, e, v, mod, diags, post, kw, states = setup
    print("spectral sums", flush=True)
    Hh = mod.H
    for beta in (0.05, 0.4):
        Z, E, E2 = ens.level_sums(e / 6, beta)
        assert Hh.partition_function(beta) == pytest.approx(Z, rel=1e-10)
        assert abs(Hh.thermal_average(beta) - E) < 1e-10
        assert abs(Hh.F_prime(beta) - (-E2 - E ** 2)) < 1e-10 * (E2 + E ** 2)
        assert abs(Hh.free_energy(beta) + np.log(Z) / beta) < 1e-9
    cfg = states["config1"]
    assert abs(np.vdot(cfg, H @ cfg).real - 2.0) < 1e-12
    want, it = ens.newton_beta(lambda b: ens.level_sums(e / 6, b), 2.0)
    assert it < 1000 and abs(mod.get_thermal_beta(cfg, 1e-10) - want) < 1e-8
    r_o, mean_o = ens.r_values(e / 6)
    r = Hh.get_r_value()
    d = np.diff(np.sort(e / 6))
    ok = np.maximum(d[:-1], d[1:]) > 1e-6
    assert ok.sum() > 0.9 * ok.size
    assert np.max(np.abs(r[ok] - r_o[ok])) < 1e-6 and abs(Hh.r_value - mean_o) < 1e-6
    X = mod.compute_expH(0.3).toarray()
    Xo = expm(-0.3 * H.toarray())
    assert np.max(np.abs(X - Xo)) < 1e-12 * np.max(np.abs(Xo))
    # partial solve: the reference's exponential branch, over the total energies
    m2 = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    m2.build_Hamiltonian(1.0, 5.0)
    m2.diagonalize_Hamiltonian(2, "sparse", tol=1e-12)
    for beta in (0.05, 0.4):
        Z, E, E2 = ens.expm_sums(H.toarray(), beta)
        assert m2.H.partition_function(beta) == pytest.approx(Z, rel=1e-10)
        assert abs(m2.H.thermal_average(beta) - E) < 1e-9
        assert abs(m2.H.F_prime(beta) - (-E2 - E ** 2)) < 1e-10 * (E2 + E ** 2)
    want, it = ens.newton_beta(lambda b: ens.expm_sums(H.toarray(), b), 2.0)
    assert it < 1000 and abs(m2.get_thermal_beta(cfg, 1e-10) - want) < 1e-8
    with pytest.raises(ValueError):
        m2.H.get_r_value()


def test_ensemble_density_entry_point(gpu):
    """edlgt_ensemble_density on a random 5 x 7 complex matrix: blocks
    {0, 1}, {2, 3, 4} of rows (4, 0, 1, 3, 2) against numpy; members outside
    the matrix and block starts that do not partition the members are
    errors."""
    import torch
    from edlgt_amd.device import ensemble_density
    print("ensemble density entry point", flush=True)
    rng = np.random.default_rng(11)
    E = rng.normal(size=(5, 7)) + 1j * rng.normal(size=(5, 7))
    rows = np.array([4, 0, 1, 3, 2])
    c = rng.normal(size=5) + 1j * rng.normal(size=5)
    got = ensemble_density(torch.as_tensor(E, device="cuda"), rows, c, [0, 2, 5]).cpu().numpy()
    want = np.abs(c[0] * E[4] + c[1] * E[0]) ** 2 + np.abs(c[2] * E[1] + c[3] * E[3] + c[4] * E[2]) ** 2
    assert np.max(np.abs(got - want)) < 1e-13
    with pytest.raises(ValueError):
        ensemble_density(torch.as_tensor(E, device="cuda"), [5], [1.0], [0, 1])
    with pytest.raises(ValueError):
        ensemble_density(torch.as_tensor(E, device="cuda"), rows, c, [0, 2, 4])
