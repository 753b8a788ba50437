"""GPU: the real-gauge solve of QED and SU(2) with staggered matter
(diagonalize_Hamiltonian(..., real_gauge=True)).

The device builds D^H H D, D = diag(i^(sum_j a_j N_j)), a_j = (x_j + z_j) mod 2,
as a float64 CSR from the same term program with one power of i per
placement, solves it with real vectors and maps the eigenvectors back through
D (edlgt_basis_quarter_phase).  Checked against the oracle: the real matrix
equals D^H H D (pattern bit-exact, values <= 1e-14 relative, the oracle's
definition pinned in tests/test_real_gauge.py), the eigenvalues equal the
complex solve's and the oracle's (<= 1e-10), the mapped vectors are
eigenvectors of the complex H (residual through the complex CSR: a wrong phase
map fails it), a non-degenerate ground state equals the complex solve's up to
a global phase, and the observables agree.  A lattice periodic along x with
an odd extent has no real gauge: the complex matrix is solved and
H.real_gauge says so.  SU(2) takes a_j = x_j mod 2 with N_tot."""
import numpy as np
import pytest

from oracle.model import QEDOracle, lowest_eigs

pytestmark = pytest.mark.gpu

CASES = [([10], [True]), ([6], [False]), ([2, 2], [False, False]), ([3, 2], [True, True]),
         ([4, 2], [False, False]), ([2, 2, 2], [True, True, True])]


def _exponents(orc, mod):
    a, name = mod._gauge_weights()
    nd = np.rint(np.real(np.einsum("sii->si", orc.ops[name]))).astype(np.int64)
    s = np.zeros(orc.N, dtype=np.int64)
    for j in range(orc.n_sites):
        s += a[j] * nd[j][orc.configs[:, j]]
    return s % 4


@pytest.mark.parametrize("lvals,obc", CASES)
def test_real_gauge_matrix_and_solve(gpu, lvals, obc):
    import torch
    from edlgt_amd import QED_Model
    print(f"real gauge: QED lvals={lvals} obc={obc}", flush=True)
    orc = QEDOracle(1, False, lvals, obc)
    mod = QED_Model(spin=1, pure_theory=False, lvals=lvals, has_obc=obc)
    assert np.array_equal(mod.sector_configs, orc.configs)
    Hc = mod.build_Hamiltonian(1.3, 0.7)
    Ho = orc.hamiltonian(1.3, 0.7)
    rg = mod.real_gauge_matrix()
    assert rg is not None
    Hr, site_exp = rg
    assert not Hr.is_complex
    # D^H H D from the oracle's complex matrix
    ph = 1j ** _exponents(orc, mod)
    C = Ho.tocoo()
    want = C.copy()
    want.data = np.conj(ph[C.row]) * C.data * ph[C.col]
    want = want.tocsr()
    want.sort_indices()
    assert np.max(np.abs(want.data.imag)) == 0.0
    rp, col, val = Hr.to_host()
    assert np.array_equal(rp, want.indptr.astype(np.int64)) and np.array_equal(col, want.indices.astype(np.int32))
    assert np.max(np.abs(val - want.data.real)) <= 1e-14 * np.max(np.abs(want.data))

    k = min(3, orc.N - 1)
    obs = dict(local_obs=["E2", "N"])
    if len(lvals) > 1:
        obs["plaquette_obs"] = [["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"]]
    mod.get_observables(**obs)
    mod.diagonalize_Hamiltonian(k, "sparse")
    assert not mod.H.real_gauge
    e_c = mod.H.energies.copy()
    psi_c = mod.H.Npsi[0].psi_device.clone()
    mod.measure_observables(0)
    res_c = {key: np.array(v, copy=True) for key, v in mod.res.items() if key != "energy"}
    mod.diagonalize_Hamiltonian(k, "sparse", real_gauge=True)
    assert mod.H.real_gauge
    e_r = mod.H.energies
    ref = lowest_eigs(Ho, k) if orc.N > 600 else np.linalg.eigvalsh(Ho.toarray())[:k]
    assert np.max(np.abs(e_r - e_c)) < 1e-10 and np.max(np.abs(e_r - ref)) < 1e-10
    for i in range(k):
        psi = mod.H.Npsi[i].psi_device
        assert psi.dtype == torch.complex128
        r = Hc.matvec(psi) - e_r[i] * psi
        assert float(torch.linalg.vector_norm(r)) <= 1e-8 * max(1.0, abs(e_r[i]))
    # QMB_state.expectation_value through the device CSR (complex H, and the
    # real matrix on the gauge-transformed vector), and through SciPy on the host
    st = mod.H.Npsi[0]
    assert abs(st.expectation_value(Hc) - e_r[0]) < 1e-9 * max(1.0, abs(e_r[0]))
    assert abs(st.expectation_value(Hc, "imag")) < 1e-9
    if orc.N <= 200_000:
        assert abs(st.expectation_value(Ho) - e_r[0]) < 1e-9 * max(1.0, abs(e_r[0]))
    assert 0.0 <= st.participation_renyi_entropy(2) <= np.log(orc.N) + 1e-12
    if e_c[1] - e_c[0] > 1e-6:  # non-degenerate ground state: same vector up to a phase
        ov = abs(complex(torch.vdot(psi_c, mod.H.Npsi[0].psi_device)))
        assert abs(ov - 1.0) < 1e-9
        mod.measure_observables(0)
        for key, v in res_c.items():
            assert np.max(np.abs(np.asarray(mod.res[key]) - v)) < 1e-9, key


def test_odd_periodic_extent_solves_the_complex_matrix(gpu):
    from edlgt_amd import QED_Model
    orc = QEDOracle(1, False, [3, 2], [False, False])
    mod = QED_Model(spin=1, pure_theory=False, lvals=[3, 2], has_obc=[False, False])
    mod.build_Hamiltonian(1.3, 0.7)
    assert mod.real_gauge_matrix() is None
    mod.diagonalize_Hamiltonian(2, "sparse", real_gauge=True)
    assert not mod.H.real_gauge
    ref = lowest_eigs(orc.hamiltonian(1.3, 0.7), 2)
    assert np.max(np.abs(mod.H.energies - ref)) < 1e-10


def test_pure_theory_has_no_real_gauge(gpu):
    from edlgt_amd import QED_Model
    mod = QED_Model(spin=1, pure_theory=True, lvals=[2, 2], has_obc=[False, False])
    mod.build_Hamiltonian(1.3)
    assert mod.real_gauge_matrix() is None


def test_real_gauge_mixed_precision_and_rebuild(gpu):
    """4x2 periodic (140,647 states): the mixed FP32/FP64 solve of the real
    matrix reaches the FP64 eigenvalue; a second build_Hamiltonian at another
    coupling replaces the cached real matrix."""
    from edlgt_amd import PREC_MIXED, QED_Model
    orc = QEDOracle(1, False, [4, 2], [False, False])
    mod = QED_Model(spin=1, pure_theory=False, lvals=[4, 2], has_obc=[False, False])
    for g in (1.0, 2.0):
        mod.build_Hamiltonian(g, 0.5)
        ref = lowest_eigs(orc.hamiltonian(g, 0.5), 1)[0]
        mod.diagonalize_Hamiltonian(1, "sparse", real_gauge=True, precision=PREC_MIXED)
        assert mod.H.real_gauge
        assert abs(mod.H.energies[0] - ref) < 1e-10, (g, mod.H.energies[0], ref)
    # the multi-term program of a coupling sweep does not replace the program
    # of the last build_Hamiltonian (g = 2) that the real matrix comes from
    mod.build_Hamiltonian(2.0, 0.5)
    mod.build_multi_term(0.5)
    assert mod.coeffs["g"] == 2.0  # the couplings of the last build are kept
    mod.diagonalize_Hamiltonian(1, "sparse", real_gauge=True)
    assert mod.H.real_gauge and abs(mod.H.energies[0] - ref) < 1e-10, (mod.H.energies[0], ref)


@pytest.mark.parametrize("lvals,obc,spin", [([6], [True], 0.5), ([6], [False], 0.5), ([2, 2], [False, False], 0.5),
                                            ([4], [True], 1.0), ([2, 2, 2], [True, True, True], 0.5)])
def test_su2_real_gauge(gpu, lvals, obc, spin):
    """SU(2) with matter (x hopping -2 sqrt(2) i, a_j = x_j mod 2, N_tot): the
    real matrix equals the oracle's D^H H D, and the real-gauge eigenpairs are
    the complex solve's (eigenvalues <= 1e-10, residual through the complex H)."""
    import torch
    from edlgt_amd import SU2_Model
    from oracle.model import SU2Oracle
    print(f"real gauge: SU(2) lvals={lvals} obc={obc} spin={spin}", flush=True)
    n = int(np.prod(lvals))
    orc = SU2Oracle(False, lvals, obc, sectors=[n], spin=spin, generic=spin > 0.5)
    mod = SU2_Model(spin=spin, pure_theory=False, sectors=[n], lvals=lvals, has_obc=obc)
    Hc = mod.build_Hamiltonian(1.3, 0.7)
    Ho = orc.hamiltonian(1.3, 0.7)
    Hr, _ = mod.real_gauge_matrix()
    a, name = mod._gauge_weights()
    nd = np.rint(np.real(np.einsum("sii->si", orc.ops[name]))).astype(np.int64)
    s = np.zeros(orc.N, dtype=np.int64)
    for j in range(n):
        s += a[j] * nd[j][orc.configs[:, j]]
    ph = 1j ** (s % 4)
    C = Ho.tocoo()
    want = C.copy()
    want.data = np.conj(ph[C.row]) * C.data * ph[C.col]
    want = want.tocsr()
    want.sort_indices()
    rp, col, val = Hr.to_host()
    assert np.array_equal(rp, want.indptr.astype(np.int64)) and np.array_equal(col, want.indices.astype(np.int32))
    assert np.max(np.abs(val - want.data.real)) <= 1e-14 * np.max(np.abs(want.data))
    k = min(3, orc.N - 1)
    mod.diagonalize_Hamiltonian(k, "sparse", real_gauge=True)
    assert mod.H.real_gauge
    ref = lowest_eigs(Ho, k) if orc.N > 600 else np.linalg.eigvalsh(Ho.toarray())[:k]
    assert np.max(np.abs(mod.H.energies - ref)) < 1e-10
    for i in range(k):
        psi = mod.H.Npsi[i].psi_device
        r = Hc.matvec(psi) - mod.H.energies[i] * psi
        assert float(torch.linalg.vector_norm(r)) <= 1e-8 * max(1.0, abs(mod.H.energies[i]))


def test_su2_odd_periodic_chain_solves_the_complex_matrix(gpu):
    from edlgt_amd import SU2_Model
    from oracle.model import SU2Oracle
    orc = SU2Oracle(False, [5], [False], sectors=[4])
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[4], lvals=[5], has_obc=[False])
    mod.build_Hamiltonian(1.3, 0.7)
    assert mod.real_gauge_matrix() is None
    mod.diagonalize_Hamiltonian(2, "sparse", real_gauge=True)
    assert not mod.H.real_gauge
    assert np.max(np.abs(mod.H.energies - np.linalg.eigvalsh(orc.hamiltonian(1.3, 0.7).toarray())[:2])) < 1e-10


@pytest.mark.parametrize("lvals,obc", [([10], [True]), ([2, 2], [False, False]), ([6], [False])])
def test_real_gauge_time_evolution(gpu, lvals, obc):
    """exp(-iHt) psi0 = D exp(-i D^H H D t) D^H psi0: the real-gauge Krylov
    states equal the complex run's (<= 1e-10) and the oracle's expm_multiply
    path (<= 1e-8), with the same per-step site marginals, from a random
    complex start state (QED with matter)."""
    from edlgt_amd import QED_Model
    from oracle.model import expm_path
    print(f"real-gauge dynamics: QED lvals={lvals} obc={obc}", flush=True)
    orc = QEDOracle(1, False, lvals, obc)
    mod = QED_Model(spin=1, pure_theory=False, lvals=lvals, has_obc=obc)
    mod.build_Hamiltonian(1.3, 0.7)
    rng = np.random.default_rng(3)
    psi0 = rng.normal(size=orc.N) + 1j * rng.normal(size=orc.N)
    psi0 /= np.linalg.norm(psi0)
    tl = np.linspace(0.0, 0.5, 6)
    mod.time_evolution_Hamiltonian(psi0, tl)
    assert not mod.H.real_gauge
    ref = [st.psi.copy() for st in mod.H.psi_time]
    marg = mod.H.marginals.copy()
    mod.time_evolution_Hamiltonian(psi0, tl, real_gauge=True)
    assert mod.H.real_gauge
    got = [st.psi for st in mod.H.psi_time]
    assert max(np.max(np.abs(a - b)) for a, b in zip(got, ref)) < 1e-10
    assert np.max(np.abs(mod.H.marginals - marg)) < 1e-10
    assert np.max(np.abs(mod.H.psi_final.cpu().numpy() - ref[-1])) < 1e-10
    exact = expm_path(orc.hamiltonian(1.3, 0.7), psi0, 0.0, 0.5, 6)
    assert max(np.max(np.abs(got[k] - exact[k])) for k in range(6)) < 1e-8
    mod.H.release_krylov()


def test_real_gauge_multi_term_sweep(gpu):
    """build_multi_term(real_gauge=True): the parts of D^H H(g) D on one real
    pattern; one server solves a coupling sweep on it with the eigenvalues of
    the complex per-coupling builds (<= 1e-10).  A lattice without a real gauge
    falls back to the complex multi-term matrix."""
    from edlgt_amd import QED_Model, Server
    mod = QED_Model(spin=1, pure_theory=False, lvals=[2, 2], has_obc=[False, False])
    gs = [0.5, 1.0, 2.0]
    plain = {}
    for g in gs:
        H = mod.build_Hamiltonian(g, 0.5)
        plain[g] = H.lanczos(4, tol=1e-12, want_vectors=False)[0]
    H3 = mod.build_multi_term(0.5, real_gauge=True)
    assert H3.real_gauge and not H3.is_complex
    res = Server(H3, n_eigs=4, tol=1e-12, max_basis=64, max_slots=3).solve_all(
        [(i, mod.multi_term_coefficients(g), 7 + i) for i, g in enumerate(gs)])
    for i, g in enumerate(gs):
        assert np.max(np.abs(res[i][0] - plain[g])) < 1e-10, (g, res[i][0], plain[g])
    odd = QED_Model(spin=1, pure_theory=False, lvals=[3, 2], has_obc=[False, False])
    H3o = odd.build_multi_term(0.5, real_gauge=True)
    assert not H3o.real_gauge and H3o.is_complex
