"""CPU: the ensemble averages restated from the reference
(oracle/ensembles.py, loops over eigenstates as in quantum_model.py:912-1156),
pinned by closed forms on the 6-site periodic SU(2) chain: the diagonal
ensemble of an eigenstate is that eigenstate's value; at beta = 0 the
canonical average is the infinite-temperature mean Tr(O) / (N L) over the
sector, at large beta the ground state's value; the microcanonical shell of a
two-eigenstate superposition holds exactly the eigenvalues strictly between
the two, and an empty shell is an error."""
import numpy as np
import pytest

from oracle import ensembles as ens
from oracle.model import SU2Oracle


@pytest.fixture(scope="module")
def chain():
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    H = orc.hamiltonian(1.0, 5.0).toarray()
    e, v = np.linalg.eigh(H)
    diag = np.real(np.einsum("sii->si", orc.ops["N_single"]))
    return orc, H, e, v, diag


def test_diagonal_ensemble_of_an_eigenstate(chain):
    """A state inside one energy level is its own diagonal ensemble: a
    non-degenerate eigenstate, and a mixture of two degenerate partners (block
    branch: the projection on the level returns the state itself)."""
    orc, H, e, v, diag = chain
    gap = np.diff(e)
    post = {"N_single": np.mean}
    k = next(i for i in range(1, len(e) - 1) if gap[i - 1] > 1e-6 and gap[i] > 1e-6)
    out = ens.diagonal(v, e / 6, v[:, k], orc.configs, {"N_single": diag}, post)
    assert abs(out["DE_N_single"] - np.mean(ens.site_values(v[:, k], orc.configs, diag))) < 1e-12
    j = int(np.nonzero(gap < 1e-10)[0][0])
    psi = 0.6 * v[:, j] + 0.8j * v[:, j + 1]
    out = ens.diagonal(v, e / 6, psi, orc.configs, {"N_single": diag}, post)
    assert abs(out["DE_N_single"] - np.mean(ens.site_values(psi, orc.configs, diag))) < 1e-12


def test_canonical_limits(chain):
    orc, H, e, v, diag = chain
    inf_t = float(np.mean(diag[np.arange(6)[None, :], orc.configs].sum(axis=1))) / 6
    assert abs(ens.canonical(e, v, orc.configs, diag, 0.0, 6) - inf_t) < 1e-12
    assert e[1] - e[0] > 0.1
    gs = float(np.mean(ens.site_values(v[:, 0], orc.configs, diag)))
    assert abs(ens.canonical(e, v, orc.configs, diag, 400.0 / (e[1] - e[0]), 6) - gs) < 1e-12


def test_microcanonical_shell(chain):
    """psi = sqrt(0.8) E_a + sqrt(0.2) E_b: <H> = e_a + 0.2 D, sigma = 0.4 D
    (D = e_b - e_a), so the shell is the open interval (e_a - 0.2 D, e_a + 0.6 D);
    a and b are chosen so that no eigenvalue lies near either end."""
    orc, H, e, v, diag = chain
    a = 10
    for b in range(a + 20, len(e)):
        D = e[b] - e[a]
        if min(np.min(np.abs(e - (e[a] - 0.2 * D))), np.min(np.abs(e - (e[a] + 0.6 * D)))) > 1e-4:
            break
    psi = np.sqrt(0.8) * v[:, a] + np.sqrt(0.2) * v[:, b]
    _, me = ens.microcanonical(H, v, e / 6, psi, orc.configs, {"N_single": diag}, {"N_single": np.mean}, 6)
    inside = np.nonzero((e > e[a] - 0.2 * D) & (e < e[a] + 0.6 * D))[0]
    assert a in inside and b not in inside
    want = np.mean([np.mean(ens.site_values(v[:, i], orc.configs, diag)) for i in inside])
    assert abs(me["ME_N_single"] - want) < 1e-12
    with pytest.raises(ValueError):
        ens.microcanonical(H, v, e / 6, np.zeros(len(e)), orc.configs, {"N_single": diag}, {"N_single": np.mean}, 6)


@pytest.mark.parametrize("beta", [0.0, 0.05, 0.4])
def test_level_sums_are_the_exponential_traces(chain, beta):
    """Tr f(H) = sum_i f(E_i): the sums over the total energies equal the
    dense-exponential traces (the reference's other branch); at beta = 0,
    Z = N and <H> = Tr(H) / N."""
    orc, H, e, v, diag = chain
    a = ens.expm_sums(H, beta)
    b = ens.level_sums(e, beta)
    for x, y in zip(a, b):
        assert abs(x - y) <= 1e-11 * max(1.0, abs(x))
    if beta == 0.0:
        assert a[0] == pytest.approx(len(e), rel=1e-14) and a[1] == pytest.approx(np.trace(H) / len(e), rel=1e-13)


def test_newton_beta_reaches_the_state_energy(chain):
    """Where the search converges, <e>_beta is the target (F = 0 up to the stop
    rule's |F| < threshold |F'|): over the
    total energies for a configuration of energy 2 (beta > 0, below the
    infinite-temperature mean 6) and of energy 10 (beta < 0)."""
    orc, H, e, v, diag = chain
    for target, sign in ((2.0, 1), (10.0, -1)):
        beta, it = ens.newton_beta(lambda b: ens.level_sums(e, b), target)
        assert it < 1000 and np.sign(beta) == sign
        _, E, E2 = ens.level_sums(e, beta)
        # the loop stops on a step below 1e-10, i.e. |F| < 1e-10 |F'|
        assert abs(E - target) < 1e-10 * (E2 + E ** 2) * 1.01


def test_r_values_closed_forms():
    r, mean = ens.r_values(np.arange(12.0) ** 1)  # equal spacing: r = 1
    assert np.all(r == 1.0) and mean == 1.0
    r, _ = ens.r_values(np.array([3.0, 0.0, 1.0, 7.0]))  # gaps 1, 2, 4
    assert np.allclose(r, [0.5, 0.5])
