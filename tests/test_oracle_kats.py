"""Pins the CPU oracle against the reference's own known-answer tests
(validation/qed/test_qed_1d_matter.py:8-72, test_qed_2d_matter.py:8-74,
test_qed_2d_pure.py:8-71: per-site ground-state energies at atol 1e-10) and
the independent transfer-matrix sector counts."""
import numpy as np
import pytest

from oracle.model import QEDOracle, lowest_eigs

KATS = [
    # lvals, has_obc, pure, g, m, E0 per site, sector dim
    ([4], [True], False, 1.0, 1.0, -0.5691411466075982, 6),
    ([4], [False], False, 1.0, 1.0, -0.5903881075118848, 13),
    ([4, 2], [True, True], False, 1.0, 1.0, -0.6862839217038432, None),
    ([3, 3], [True, True], True, 1.0, None, -0.09443358823578024, None),
    ([3, 3], [False, False], True, 1.0, None, -0.2042469295595554, None),
]


@pytest.mark.parametrize("lvals,obc,pure,g,m,e0,dim", KATS)
def test_qed_kat(lvals, obc, pure, g, m, e0, dim):
    M = QEDOracle(1, pure, lvals, obc)
    if dim is not None:
        assert M.N == dim
    H = M.hamiltonian(g, m)
    e = lowest_eigs(H, 1)[0] / np.prod(lvals)
    assert abs(e - e0) < 1e-10


def test_qed_2d_pbc_kat_and_c2_dimension():
    M = QEDOracle(1, False, [4, 2], [False, False])
    assert M.N == 140647
    H = M.hamiltonian(1.0, 1.0)
    e = lowest_eigs(H, 1)[0] / 8
    assert abs(e - (-0.8830087027925408)) < 1e-10


def test_c1_dimension():
    assert QEDOracle(1, False, [10], [True]).N == 197


def test_qed_3d_theta_kat():
    """validation/qed/test_qed_3d_theta.py: 2x2x2 pure PBC (flux through the
    three coordinate planes at the origin fixed to 0), g=2.3, theta=0.41."""
    M = QEDOracle(1, True, [2, 2, 2], [False, False, False])
    e = lowest_eigs(M.hamiltonian(2.3, theta=0.41), 1)[0] / 8
    assert abs(e - (-0.37555632787741716)) < 1e-10


def _qed_ground_observables(M, psi):
    """workflows/qed.py:134-155 on the oracle: E2 = link_avg('E2') over the
    existing links, N = 0.5 even(N) + 0.5 odd(N_zero), mean xy plaquette."""
    from oracle import lattice as lat
    from oracle.model import expval_nbody
    p2 = np.abs(psi) ** 2
    cfg = M.configs

    def local(name):
        return np.array([np.sum(p2 * np.real(M.ops[name][s][cfg[:, s], cfg[:, s]])) for s in range(M.n_sites)])

    e2, cnt = 0.0, 0
    for d in M.dirs:
        v = local(f"E2_p{d}")
        for i in range(M.n_sites):
            if lat.neighbor_sites(lat.zig_zag(M.lvals, i), M.lvals, d, M.has_obc) is not None:
                e2 += v[i]
                cnt += 1
    even = np.array([lat.staggered_mask(M.lvals, "even")[lat.zig_zag(M.lvals, i)] for i in range(M.n_sites)])
    n_obs = 0.5 * local("N")[even].mean() + 0.5 * local("N_zero")[~even].mean()
    names = ["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"]
    plaqs = [p for p in (lat.plaquette_sites(lat.zig_zag(M.lvals, i), M.lvals, ["x", "y"], M.has_obc)
                         for i in range(M.n_sites)) if p]
    plq = np.mean([expval_nbody([M.ops[n] for n in names], p, cfg, M.keys, M.loc_dims, psi).real for p in plaqs])
    return e2 / cnt, n_obs, plq


def test_qed_2d_stringbreaking_kat():
    """validation/qed/test_qed_2d_stringbreaking.py:8-36: 3x2 OBC with matter
    and static charges -1 (site 0) and +1 (site 5), g = m = 1: energy per
    site, E2, N and the plaquette at atol 1e-10."""
    M = QEDOracle(1, False, [3, 2], [True, True], bg_list=[-1, 0, 0, 0, 0, 1])
    H = M.hamiltonian(1.0, 1.0)
    w, v = np.linalg.eigh(H.toarray())
    assert abs(w[0] / 6 - (-0.49619430243949)) < 1e-10
    e2, n_obs, plq = _qed_ground_observables(M, v[:, 0])
    assert abs(e2 - 0.4462942942081073) < 1e-10
    assert abs(n_obs - 0.08360798337722186) < 1e-10
    assert abs(plq - 0.4555429224294397) < 1e-10


def test_background_charges_shift_gauss_law():
    """With static charges the kept local states satisfy div E = matter
    charge + q (QED_operators.py:503-510), the operator 'bg' is q on them, and
    a zero background is the plain model."""
    M = QEDOracle(1, False, [3, 2], [True, True], bg_list=[-1, 0, 0, 0, 0, 1])
    M0 = QEDOracle(1, False, [3, 2], [True, True])
    Mz = QEDOracle(1, False, [3, 2], [True, True], bg_list=[0] * 6)
    assert np.array_equal(Mz.configs, M0.configs)
    for s, q in enumerate([-1, 0, 0, 0, 0, 1]):
        d = M.loc_dims[s]
        assert np.allclose(M.ops["bg"][s][:d, :d], q * np.eye(d))
        div = sum(np.real(np.diag(M.ops[f"E_p{a}"][s] - M.ops[f"E_m{a}"][s]))[:d] for a in M.dirs)
        odd = 1 if s in (1, 3, 5) else 0  # 3x2: x fastest, (-1)^(x+y)
        charge = np.real(np.diag(M.ops["N"][s]))[:d] - odd + q
        assert np.allclose(div, charge)
    assert M.N != M0.N


def test_label_configs_are_vacuum_and_meson():
    """The oracle's benchmark product states lie in the sector and carry the
    stated physics: the vacuum has no electric energy and the staggered
    filling; the meson adds one unit of flux on one link."""
    M = QEDOracle(1, False, [10], [True])
    rows = {tuple(r) for r in M.configs.tolist()}
    v, mc = M.label_config("V"), M.label_config("meson_center")
    assert tuple(v) in rows and tuple(mc) in rows
    e2 = lambda c: sum(np.real(M.ops["E2"][s][c[s], c[s]]) for s in range(10))
    n = lambda c: [np.real(M.ops["N"][s][c[s], c[s]]) for s in range(10)]
    assert e2(v) == 0.0 and n(v) == [0.0, 1.0] * 5
    assert e2(mc) == 1.0  # two half-links at E = +1, E2 = 1/2 sum of halves
    assert np.sum(np.array(n(mc)) != np.array(n(v))) == 2


def test_integrated_qed_oracle_pins():
    """oracle/integrated.py: free fermions at g = 0 (sum of the L/2 lowest
    one-particle energies) and Gauss's law on product states."""
    from oracle import integrated as iq
    for L, m in ((4, 0.7), (6, 0.3), (8, 1.1)):
        H = iq.hamiltonian(L, 0.0, m, np.zeros(L))
        assert abs(np.linalg.eigvalsh(H.toarray())[0] - iq.free_fermion_e0(L, m)) < 1e-12
        assert np.max(np.abs((H - H.conj().T).toarray())) < 1e-15
    L, g, m = 6, 1.3, 0.4
    cfg = iq.sector_configs(L)
    vac = np.array([1 if i % 2 == 0 else 0 for i in range(L)])
    meson = vac.copy()
    meson[2], meson[3] = 0, 1  # central (even, odd) pair flipped
    rows = {tuple(c): r for r, c in enumerate(cfg.tolist())}
    mass = lambda c: sum((m if i % 2 == 0 else -m) * (c[i] == 0) for i in range(L))
    Hg = iq.hamiltonian(L, g, m, np.zeros(L)).toarray()
    assert abs(Hg[rows[tuple(vac)], rows[tuple(vac)]] - mass(vac)) < 1e-13  # no flux
    # meson: one unit of flux on the link between its two sites
    assert abs(Hg[rows[tuple(meson)], rows[tuple(meson)]] - (mass(meson) + g / 2)) < 1e-13
    # a static charge +1 on site 0 and -1 on the last site: flux 1 on every link
    q = np.zeros(L)
    q[0], q[-1] = 1.0, -1.0
    Hq = iq.hamiltonian(L, g, m, q).toarray()
    assert abs(Hq[rows[tuple(vac)], rows[tuple(vac)]] - (mass(vac) + g / 2 * (L - 1))) < 1e-13
    psi = np.zeros(len(cfg))
    psi[rows[tuple(meson)]] = 1.0
    assert np.allclose(iq.link_e2(L, np.zeros(L), psi), [0, 0, 1, 0, 0])


def test_link_penalty_spectrum_equals_link_sector():
    """QED_Model(link_symmetries=False) (QED_model.py:131-141, 381-393): the
    sector without link constraints plus 100 E2 - 100 E_p E_m per link (=
    50 (E_p - E_m)^2 per link, commuting with H) has the link-symmetric
    sector's spectrum at the bottom, exactly."""
    from oracle.model import QEDOracle, lowest_eigs
    for lv, obc, pure, g, m in (([4], [True], False, 1.0, 1.0), ([4], [False], False, 1.0, 1.0),
                                ([2, 2], [True, True], True, 1.0, None), ([2, 2], [False, False], True, 0.7, None)):
        a = QEDOracle(1, pure, lv, obc)
        b = QEDOracle(1, pure, lv, obc, link_symmetries=False)
        assert b.N > a.N
        k = min(3, a.N - 1)
        ea = lowest_eigs(a.hamiltonian(g, m), k)
        eb = lowest_eigs(b.hamiltonian(g, m), k)
        assert np.max(np.abs(ea - eb)) < 1e-10, (lv, obc, ea, eb)
