"""SU(2) (spin-1/2 rishons): the oracle against the reference's known answers
(validation/su2/test_su2_spectrum.py: L=6 PBC, N_tot=6, g=1, m=5) and golden
dynamics (validation/su2/test_su2_dynamics.py, fixture copied from
gold_dynamics_case01_dense.npz), against brute force, and against the
mathematics of the singlet space; the engine's dressed sites against the
oracle's (bases and projected operators)."""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import su2_ops
from oracle.model import SU2Oracle, encode, expm_path, lowest_eigs

GOLD = Path(__file__).parent / "golden" / "su2_dynamics_L6_pbc.npz"


def _site_obs(orc, psi, names):
    """<O_s> per site of diagonal local operators from |psi|^2."""
    p = np.abs(psi) ** 2
    out = {}
    for n in names:
        d = np.real(np.einsum("sii->si", orc.ops[n]))
        out[n] = np.array([np.sum(p * d[s, orc.configs[:, s]]) for s in range(orc.n_sites)])
    return out


def _workflow_obs(orc, psi):
    """su2_measure_on_states aggregation (workflows/su2.py:205-221)."""
    o = _site_obs(orc, psi, ["T2_px", "N_single", "N_pair", "N_zero"])
    even = np.array([i % 2 == 0 for i in range(orc.n_sites)])
    res = {"E2": o["T2_px"].mean(), "N_single": o["N_single"].mean()}
    res["N_pair"] = 0.5 * o["N_pair"][even].mean() + 0.5 * o["N_zero"][~even].mean()
    res["N_zero"] = 0.5 * o["N_zero"][even].mean() + 0.5 * o["N_pair"][~even].mean()
    res["N_tot"] = res["N_single"] + 2 * res["N_pair"]
    return res


def _entropy(configs, psi, keep):
    """Bipartite von Neumann entropy, base 2, eigenvalues > 1e-10
    (qmb_state.py:245-291)."""
    env = [i for i in range(configs.shape[1]) if i not in keep]
    _, ri = np.unique(configs[:, keep], axis=0, return_inverse=True)
    _, ci = np.unique(configs[:, env], axis=0, return_inverse=True)
    M = np.zeros((ri.max() + 1, ci.max() + 1), dtype=complex)
    M[ri.ravel(), ci.ravel()] = psi
    lam = np.linalg.svd(M, compute_uv=False) ** 2
    lam = lam[lam > 1e-10]
    return float(-np.sum(lam * np.log2(lam)))


@pytest.fixture(scope="module")
def kat():
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    H = orc.hamiltonian(1.0, 5.0)
    w, v = np.linalg.eigh(H.toarray())
    return orc, H, w, v[:, 0]


def test_su2_kat_spectrum(kat):
    orc, H, w, psi = kat
    assert orc.N == 282 and list(orc.loc_dims) == [6] * 6
    assert abs(w[0] / 6 - (-6.173752796132477)) < 1e-10
    res = _workflow_obs(orc, psi)
    ref = {"E2": 0.0553604161539725, "N_tot": 0.15535989251437904, "N_single": 0.14157936434695081,
           "N_pair": 0.006890264083714107, "N_zero": 0.8515303715693359}
    for k, v in ref.items():
        assert abs(res[k] - v) < 1e-10, k
    assert abs(_entropy(orc.configs, psi, [0, 1, 2]) - 0.802957045143) < 1e-10


def test_su2_gold_dynamics(kat):
    orc, H, _, _ = kat
    g = np.load(GOLD, allow_pickle=False)
    t = g["time_steps"]
    cfg = np.array([0 if i % 2 == 0 else 4 for i in range(6)])
    idx = np.searchsorted(orc.keys, encode(cfg[None], orc.loc_dims))[0]
    assert np.array_equal(orc.configs[idx], cfg)
    psi0 = np.zeros(orc.N, dtype=complex)
    psi0[idx] = 1.0
    path = expm_path(H, psi0, t[0], t[-1], len(t))
    F = np.abs(path @ psi0.conj()) ** 2
    E2 = np.array([_workflow_obs(orc, p)["E2"] for p in path])
    S = np.array([_entropy(orc.configs, p, [0, 1, 2]) for p in path])
    for got, ref in ((F, g["F"]), (E2, g["E2"]), (S, g["S"])):
        d = got - ref
        assert np.max(np.abs(d)) < 1e-6 and np.sqrt(np.mean(d ** 2)) < 1e-6


def test_su2_sector_hamiltonian_equals_bruteforce():
    """L=4 PBC: the sector Hamiltonian equals the Kronecker-product Hamiltonian
    over the projected local spaces restricted to the sector."""
    orc = SU2Oracle(False, [4], [False], sectors=[4])
    g, m = 0.7, 1.3
    d = 6
    c = orc.couplings(g, m)

    def loc(name, s):
        return sp.csr_matrix(orc.ops[name][s][:d, :d])

    def place(mats):
        out = sp.identity(1, format="csr")
        for s in range(4):
            out = sp.kron(out, mats.get(s, sp.identity(d)), format="csr")
        return out

    H = sum(c["E"] * place({s: loc("E_square", s)}) for s in range(4))
    for s in range(4):
        H = H + c["m_even" if s % 2 == 0 else "m_odd"] * place({s: loc("N-1", s)})
    for s in range(4):
        j = (s + 1) % 4
        t = c["tx_even" if s % 2 == 0 else "tx_odd"]
        hop = t * (place({s: loc("Qpx_dag", s)}) @ place({j: loc("Qmx", j)}))
        H = H + hop + hop.conj().T
    idx = encode(orc.configs, orc.loc_dims)
    Hb = H.tocsr()[idx][:, idx].toarray()
    Hs = orc.hamiltonian(g, m).toarray()
    assert np.max(np.abs(Hs - Hb)) < 1e-13


@pytest.mark.parametrize("pure,dim", [(False, 1), (True, 1), (True, 2), (False, 2)])
def test_singlet_space_is_the_gauss_kernel(pure, dim):
    """The site basis is an isometry onto the common kernel of the total
    SU(2) generators (the Gauss law), with the full kernel dimension."""
    basis = su2_ops.gauge_invariant_states(pure, dim)
    B = basis["site"].toarray()
    T = su2_ops.total_spin_operators(pure, dim)
    assert np.allclose(B.T @ B, np.eye(B.shape[1]), atol=1e-14)
    for a in "xyz":
        assert np.max(np.abs(T[a] @ B)) < 1e-13
    C = sum((T[a].conj().T @ T[a]).toarray() for a in "xyz")
    kernel = int(np.sum(np.linalg.eigvalsh(C) < 1e-10))
    assert kernel == B.shape[1]
    assert B.shape[1] == {(False, 1): 6, (True, 1): 2, (True, 2): 9, (False, 2): 30}[(pure, dim)]


@pytest.mark.parametrize("pure,dim", [(False, 1), (True, 1), (True, 2), (False, 2)])
def test_engine_site_matches_oracle(pure, dim):
    from edlgt_amd.su2_operators import SU2Site
    site = SU2Site(pure, dim)
    basis = su2_ops.gauge_invariant_states(pure, dim)
    ops = su2_ops.dressed_site_operators(pure, dim)
    assert set(site.basis) == set(basis)
    for lab, Bo in basis.items():
        Bo = Bo.toarray()
        assert site.basis[lab].shape == Bo.shape
        assert np.max(np.abs(site.basis[lab] - Bo)) < 1e-15
        mine = site.operators(lab)
        for name, O in ops.items():
            ref = Bo.T @ O.toarray() @ Bo
            assert np.max(np.abs(mine[name] - ref)) < 1e-14, (lab, name)


def test_cg_against_sympy():
    from fractions import Fraction
    from math import sqrt

    from sympy import Rational
    from sympy.physics.wigner import clebsch_gordan

    from edlgt_amd.su2_operators import cg_exact
    halves = [Fraction(k, 2) for k in range(0, 5)]
    for j1 in halves:
        for j2 in halves:
            for J in halves:
                for m1 in [j1 - k for k in range(int(2 * j1) + 1)]:
                    for m2 in [j2 - k for k in range(int(2 * j2) + 1)]:
                        M = m1 + m2
                        s, r = cg_exact(j1, m1, j2, m2, J, M)
                        mine = float(s) * sqrt(float(r))
                        R = lambda f: Rational(f.numerator, f.denominator)
                        ref = float(clebsch_gordan(R(j1), R(j2), R(J), R(m1), R(m2), R(M)))
                        assert abs(mine - ref) < 1e-15


def test_su2_sector_counts():
    assert SU2Oracle(False, [10], [True], sectors=[10]).N == 8953
    assert SU2Oracle(True, [3, 2], [False, False]).N > 0


def test_expval_oracle_equals_bruteforce():
    """The oracle's n-body expectation values (sector triplets + exp_val_data)
    equal psi^H (O_i x O_j) psi over the full Kronecker space, L=4 PBC, for a
    diagonal two-body pair at every distance and the nearest-neighbour hopping."""
    from oracle.model import expval_nbody, twobody_corr
    orc = SU2Oracle(False, [4], [False], sectors=[4])
    d = 6
    rng = np.random.default_rng(4)
    psi = rng.standard_normal(orc.N) + 1j * rng.standard_normal(orc.N)
    psi /= np.linalg.norm(psi)
    full = np.zeros(d ** 4, dtype=complex)
    full[encode(orc.configs, orc.loc_dims)] = psi

    def place(mats):
        out = sp.identity(1, format="csr")
        for s in range(4):
            out = sp.kron(out, sp.csr_matrix(mats.get(s, np.eye(d))), format="csr")
        return out

    corr = twobody_corr(orc.ops["N_tot"], orc.ops["N_single"], orc.configs, orc.keys, orc.loc_dims, psi)
    for i in range(4):
        assert corr[i, i] == 0.0
        for j in range(i + 1, 4):
            O = place({i: orc.ops["N_tot"][i][:d, :d], j: orc.ops["N_single"][j][:d, :d]})
            ref = np.vdot(full, O @ full).real
            assert abs(corr[i, j] - ref) < 1e-13 and corr[j, i] == corr[i, j]
    for i in range(4):
        j = (i + 1) % 4
        got = expval_nbody([orc.ops["Qpx_dag"], orc.ops["Qmx"]], [i, j], orc.configs, orc.keys, orc.loc_dims, psi)
        O = place({i: orc.ops["Qpx_dag"][i][:d, :d]}) @ place({j: orc.ops["Qmx"][j][:d, :d]})
        assert abs(got - np.vdot(full, O @ full)) < 1e-13


def test_blockwise_schmidt_weights_equal_dense_svd():
    """The oracle's block-by-block Schmidt decomposition (connected components
    of the subsystem/environment pairs) equals one dense SVD, L=10 OBC."""
    from oracle.entropy import block_schmidt_weights as schmidt_weights
    orc = SU2Oracle(False, [10], [True], sectors=[10])
    rng = np.random.default_rng(9)
    psi = rng.standard_normal(orc.N) + 1j * rng.standard_normal(orc.N)
    psi /= np.linalg.norm(psi)
    for keep in ([0, 1, 2, 3, 4], [0, 2, 5], [9]):
        env = [i for i in range(10) if i not in keep]
        _, a = np.unique(orc.configs[:, keep], axis=0, return_inverse=True)
        _, e = np.unique(orc.configs[:, env], axis=0, return_inverse=True)
        a, e = a.ravel(), e.ravel()
        part = dict(subsys_map=a, env_map=e, subsys_dim=int(a.max()) + 1, env_dim=int(e.max()) + 1)
        M = np.zeros((part["subsys_dim"], part["env_dim"]), dtype=complex)
        M[a, e] = psi
        ref = np.sort(np.linalg.svd(M, compute_uv=False) ** 2)[::-1]
        got = np.sort(schmidt_weights(psi, part))[::-1]
        n = min(ref.size, got.size)
        assert np.max(np.abs(got[:n] - ref[:n])) < 1e-13
        assert np.all(ref[n:] < 1e-13) and np.all(got[n:] < 1e-13)
        assert abs(np.sum(got) - 1) < 1e-12


# ---------------------------------------------------------------- generalised and 3D operator sets
@pytest.mark.parametrize("j2", [1, 2, 3])
def test_generalised_rishon_algebra(j2):
    """SU2_Rishon_gen's own checks (SU2_rishons.py:179-194): the rishons
    transform as a spin-1/2 doublet under the generators and anticommute with
    the parity, for rishon truncations j_max = 1/2, 1, 3/2."""
    from sympy import Rational
    o = su2_ops.rishon_gen(Rational(j2, 2))
    comm = lambda a, b: (a @ b - b @ a).toarray()
    Zr, Zg = o["Zr"], o["Zg"]
    assert np.allclose(2 * comm(Zr, o["Tx"]), Zg.toarray(), atol=1e-14)
    assert np.allclose(2 * comm(Zg, o["Tx"]), Zr.toarray(), atol=1e-14)
    assert np.allclose(comm(Zr, o["Ty"]), -0.5j * Zg.toarray(), atol=1e-14)
    assert np.allclose(comm(Zg, o["Ty"]), 0.5j * Zr.toarray(), atol=1e-14)
    assert np.allclose(2 * comm(Zr, o["Tz"]), Zr.toarray(), atol=1e-14)
    assert np.allclose(2 * comm(Zg, o["Tz"]), -Zg.toarray(), atol=1e-14)
    for c in "rg":
        Z = o[f"Z{c}"]
        assert np.max(np.abs((Z @ o["P"] + o["P"] @ Z).toarray())) < 1e-15
    # the engine's dense rishons and generators are the same matrices
    from fractions import Fraction

    from edlgt_amd.su2_operators import rishon_matrices_gen
    mine = rishon_matrices_gen(Fraction(j2, 2))
    for name in ("Zr", "Zg", "ZA_r", "ZA_g", "ZA_r_P", "P_ZB_g_dag", "T2", "Tz", "Tx", "Ty", "P"):
        assert np.max(np.abs(mine[name] - o[name].toarray())) < 1e-15, name


@pytest.mark.parametrize("pure,dim,j2", [(False, 1, 2), (True, 2, 2), (False, 1, 3), (True, 3, 1)])
def test_generalised_singlet_space_is_the_gauss_kernel(pure, dim, j2):
    from sympy import Rational
    j = Rational(j2, 2)
    basis = su2_ops.gauge_invariant_states(pure, dim, j)
    B = basis["site"].toarray()
    T = su2_ops.total_spin_operators(pure, dim, j)
    assert np.allclose(B.T @ B, np.eye(B.shape[1]), atol=1e-14)
    for a in "xyz":
        assert np.max(np.abs(T[a] @ B)) < 1e-13
    C = sum((T[a].conj().T @ T[a]).toarray() for a in "xyz")
    assert int(np.sum(np.linalg.eigvalsh(C) < 1e-10)) == B.shape[1]


@pytest.mark.parametrize("pure,dim,j2,generic", [(False, 1, 1, True), (True, 2, 1, True), (False, 2, 1, True),
                                                  (False, 1, 2, True), (True, 2, 2, True), (True, 3, 1, False),
                                                  (False, 3, 1, False)])
def test_engine_site_matches_oracle_general(pure, dim, j2, generic):
    """Bases and every projected operator of the engine's site against the
    oracle's, for the generalised set and the 3D hard-coded set."""
    from fractions import Fraction

    from sympy import Rational

    from edlgt_amd.su2_operators import SU2Site
    site = SU2Site(pure, dim, Fraction(j2, 2), generic)
    j = Rational(j2, 2)
    basis = su2_ops.gauge_invariant_states(pure, dim, j)
    ops = su2_ops.dressed_site_operators_gen(pure, dim, j) if generic else su2_ops.dressed_site_operators(pure, dim)
    assert set(site.basis) == set(basis)
    for lab in ("site", sorted(basis)[-1]):
        Bo = basis[lab].toarray()
        assert np.max(np.abs(site.basis[lab] - Bo)) < 1e-15
        mine = site.operators(lab)
        assert set(mine) == set(ops)
        for name, O in ops.items():
            ref = Bo.T @ (O @ Bo)
            assert np.max(np.abs(mine[name] - ref)) < 1e-14, (lab, name)


@pytest.mark.parametrize("lvals,obc,pure", [([4], [True], False), ([4], [False], False),
                                            ([2, 2], [True, True], False), ([2, 2], [False, False], False),
                                            ([2, 2], [True, True], True), ([2, 2], [False, False], True),
                                            ([2, 2, 2], [True] * 3, True)])
def test_generalised_equals_hard_coded_at_spin_half(lvals, obc, pure):
    """validation/su2/test_su2_hamiltonians.py (compare_SU2_models,
    workflows/su2.py:36-89): the generalised operator set at spin 1/2 gives
    the hard-coded model's sector and its lowest eigenvalues (g = m = 0.1,
    atol 1e-10)."""
    from oracle.model import lowest_eigs
    n = int(np.prod(lvals))
    sec = None if pure else [n]
    m = None if pure else 0.1
    a = SU2Oracle(pure, lvals, obc, sectors=sec)
    b = SU2Oracle(pure, lvals, obc, sectors=sec, generic=True)
    assert np.array_equal(a.configs, b.configs)
    k = min(4 if len(lvals) <= 2 else 2, a.N - 1)
    if k < 1:
        return
    ea = np.linalg.eigvalsh(a.hamiltonian(0.1, m).toarray())[:k] if a.N <= 600 else lowest_eigs(a.hamiltonian(0.1, m), k)
    eb = np.linalg.eigvalsh(b.hamiltonian(0.1, m).toarray())[:k] if b.N <= 600 else lowest_eigs(b.hamiltonian(0.1, m), k)
    assert np.max(np.abs(ea - eb)) / n < 1e-10, (ea, eb)


def test_su2_3d_theta_term_oracle_pins():
    """3D theta term (E.B per plane, SU2_model.py:303-336): H(theta) is
    Hermitian, H(-theta) = conj(H(theta)) (the term is odd under time
    reversal), so the spectrum is even in theta; and the term changes it."""
    orc = SU2Oracle(True, [2, 2, 2], [True] * 3)
    H0 = orc.hamiltonian(1.2).toarray()
    Hp = orc.hamiltonian(1.2, theta=0.7).toarray()
    Hm = orc.hamiltonian(1.2, theta=-0.7).toarray()
    assert np.max(np.abs(Hp - Hp.conj().T)) < 1e-14
    assert np.max(np.abs(Hm - Hp.conj())) < 1e-14
    ep, em, e0 = (np.linalg.eigvalsh(h) for h in (Hp, Hm, H0))
    assert np.max(np.abs(ep - em)) < 1e-12
    assert np.max(np.abs(ep - e0)) > 1e-3


def test_engine_su2_3d_theta_operators_match_oracle():
    from fractions import Fraction

    from edlgt_amd.su2_operators import SU2Site
    site = SU2Site(True, 3, Fraction(1, 2), False)
    basis = su2_ops.gauge_invariant_states(True, 3)
    ops = su2_ops.dressed_site_operators(True, 3)
    B = basis["site"].toarray()
    mine = site.operators("site")
    for name in ("EzC_px,py", "EyC_px,pz", "ExC_py,pz"):
        ref = B.T @ (ops[name] @ B)
        assert np.max(np.abs(mine[name] - ref)) < 1e-14, name
        assert np.max(np.abs(ref)) > 0.1


# ---------------------------------------------------------------- background charges
@pytest.mark.parametrize("L,qpos", [(4, (0, 3)), (6, (0, 5)), (6, (2, 3)), (5, (1, 4))])
def test_su2_background_flux_string_closed_form(L, qpos):
    """Pure 1D SU(2), OBC, spin-1/2 static charges at two sites: Gauss's law
    forces j = 1/2 on every link between them and j = 0 elsewhere, a single
    sector state whose energy is E sum_half-links T2/2 = (8g/3)(3/4) per link
    = 2g per link (SU2_model.py:555-624 couplings, E_square = sum T2/2)."""
    bg = [0.0] * L
    bg[qpos[0]] = bg[qpos[1]] = 0.5
    g = 1.3
    orc = SU2Oracle(True, [L], [True], bg_list=bg)
    assert orc.N == 1
    assert abs(orc.hamiltonian(g).toarray()[0, 0] - 2 * g * (qpos[1] - qpos[0])) < 1e-13
    # every site's background Casimir is its charge's j(j+1)
    for i in range(L):
        d = orc.loc_dims[i]
        assert np.allclose(np.diag(orc.ops["bg"][i][:d, :d]).real, bg[i] * (bg[i] + 1))


def test_su2_zero_background_is_the_plain_model():
    a = SU2Oracle(False, [4], [True], sectors=[4])
    b = SU2Oracle(False, [4], [True], sectors=[4], bg_list=[0, 0, 0, 0])
    assert b.bg_list is None and np.array_equal(a.configs, b.configs)


@pytest.mark.parametrize("pure,dim", [(False, 1), (True, 1), (True, 2), (False, 2)])
def test_engine_site_with_background_matches_oracle(pure, dim):
    from fractions import Fraction

    from edlgt_amd.su2_operators import SU2Site
    site = SU2Site(pure, dim, Fraction(1, 2), False, background=Fraction(1, 2))
    basis, bgcol = su2_ops.gauge_invariant_states(pure, dim, background=Fraction(1, 2), return_bg=True)
    ops = su2_ops.dressed_site_operators(pure, dim, background=Fraction(1, 2))
    for lab, Bo in basis.items():
        for k in (0, 1):
            cols = np.flatnonzero(np.isclose(bgcol[lab], k / 2))
            mine_lab = f"{lab}@bg{k}"
            if cols.size == 0:
                assert mine_lab not in site.basis
                continue
            B = Bo.toarray()[:, cols]
            assert np.max(np.abs(site.basis[mine_lab] - B)) < 1e-15
            mine = site.operators(mine_lab)
            assert set(mine) == set(ops), (set(mine) ^ set(ops))
            for name, O in ops.items():
                assert np.max(np.abs(mine[name] - B.T @ (O @ B))) < 1e-14, (mine_lab, name)


@pytest.mark.parametrize("pure,dim,j2,bg2", [(False, 1, 1, 1), (True, 1, 2, 2), (False, 1, 2, 1), (True, 2, 1, 1),
                                             (True, 2, 2, 1)])
def test_engine_site_with_background_matches_oracle_general(pure, dim, j2, bg2):
    """The generalised set with a background slot (SU2_operators.py:302-590):
    bases per background charge and every projected operator of the
    engine's site against the oracle's."""
    from fractions import Fraction

    from sympy import Rational

    from edlgt_amd.su2_operators import SU2Site
    site = SU2Site(pure, dim, Fraction(j2, 2), True, background=Fraction(bg2, 2))
    j, bg = Rational(j2, 2), Rational(bg2, 2)
    basis, bgcol = su2_ops.gauge_invariant_states(pure, dim, j, background=bg, return_bg=True)
    ops = su2_ops.dressed_site_operators_gen(pure, dim, j, background=bg)
    assert "bg" in ops and "Vpx_dag" not in ops
    n_checked = 0
    for lab, Bo in basis.items():
        for k in range(bg2 + 1):
            cols = np.flatnonzero(np.isclose(bgcol[lab], k / 2))
            mine_lab = f"{lab}@bg{k}"
            if cols.size == 0:
                assert mine_lab not in site.basis
                continue
            B = Bo.toarray()[:, cols]
            assert np.max(np.abs(site.basis[mine_lab] - B)) < 1e-15
            mine = site.operators(mine_lab)
            assert set(mine) == set(ops), (set(mine) ^ set(ops))
            for name, O in ops.items():
                assert np.max(np.abs(mine[name] - B.T @ (O @ B))) < 1e-14, (mine_lab, name)
            # the background Casimir is diagonal with the charge's j(j+1)
            assert np.allclose(mine["bg"], (k / 2) * (k / 2 + 1) * np.eye(B.shape[1]))
            n_checked += 1
    assert n_checked >= 2


@pytest.mark.parametrize("lvals,obc,pure,bg", [([4], [True], False, [0, 0.5, 0, 0.5]),
                                               ([4], [False], False, [0.5, 0, 0.5, 0]),
                                               ([4], [False], True, [0.5, 0, 0, 0.5]),
                                               ([2, 2], [True, True], True, [0.5, 0, 0, 0.5])])
def test_generalised_equals_hard_coded_with_background(lvals, obc, pure, bg):
    """With spin-1/2 background charges the generalised set at spin 1/2 gives
    the hard-coded model's sector and lowest eigenvalues (g = m = 0.1)."""
    n = int(np.prod(lvals))
    sec = None if pure else [n]
    m = None if pure else 0.1
    a = SU2Oracle(pure, lvals, obc, sectors=sec, bg_list=bg)
    b = SU2Oracle(pure, lvals, obc, sectors=sec, generic=True, bg_list=bg)
    assert np.array_equal(a.configs, b.configs)
    k = min(4, a.N)
    ea = np.linalg.eigvalsh(a.hamiltonian(0.1, m).toarray())[:k]
    eb = np.linalg.eigvalsh(b.hamiltonian(0.1, m).toarray())[:k]
    assert np.max(np.abs(ea - eb)) / n < 1e-10, (ea, eb)


@pytest.mark.parametrize("L,bg,links", [(4, [1, 0, 0, 1], [1, 1, 1]), (5, [0, 1, 0, 1, 0], [0, 1, 1, 0]),
                                        (4, [0.5, 0, 1, 0.5], [0.5, 0.5, 0.5]),
                                        (5, [0.5, 0, 0, 0.5, 0], [0.5, 0.5, 0.5, 0])])
def test_su2_spin1_background_flux_string_closed_form(L, bg, links):
    """Pure 1D SU(2), OBC, rishons truncated at spin 1, static charges of
    spin 1/2 and 1: Gauss's law fixes every link's spin (j on a link equals
    the coupling of the charges to its left, within the truncation), so the
    sector holds one state, of energy (8g/3) sum_links j(j+1)
    (two half-links of T2/2 each; SU2_model.py:555-624)."""
    g = 0.7
    orc = SU2Oracle(True, [L], [True], spin=1, bg_list=bg)
    assert orc.generic and orc.N == 1
    e = orc.hamiltonian(g).toarray()[0, 0]
    assert abs(e - (8 * g / 3) * sum(j * (j + 1) for j in links)) < 1e-12
