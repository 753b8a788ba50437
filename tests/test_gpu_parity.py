"""GPU parity: device sector / CSR / eigenvalues against the CPU oracle.

Bars: sector configurations and CSR pattern bit-exact; matrix values within
1e-14 relative (max |dv| <= 1e-14 * max |v|); eigenvalues within 1e-10 of the
oracle and of the reference's known-answer energies
(validation/qed/test_qed_1d_matter.py, test_qed_2d_matter.py, test_qed_2d_pure.py)."""
import numpy as np
import pytest

from oracle.model import QEDOracle, lowest_eigs

pytestmark = pytest.mark.gpu

CASES = [
    # lvals, has_obc, pure, g, m, KAT energy per site (or None)
    ([4], [True], False, 1.0, 1.0, -0.5691411466075982),
    ([4], [False], False, 1.0, 1.0, -0.5903881075118848),
    ([10], [True], False, 1.0, 1.0, None),
    ([4, 2], [True, True], False, 1.0, 1.0, -0.6862839217038432),
    ([3, 3], [True, True], True, 1.0, None, -0.09443358823578024),
    ([3, 3], [False, False], True, 1.0, None, -0.2042469295595554),
    ([2, 2], [False, False], False, 0.7, 0.3, None),
]


def _model(lvals, obc, pure):
    from edlgt_amd import QED_Model
    return QED_Model(spin=1, pure_theory=pure, lvals=lvals, has_obc=obc)


def _csr_match(Hg, Ho):
    Ho = Ho.tocsr()
    rp, col, val = Hg.to_host()
    assert np.array_equal(rp, Ho.indptr.astype(np.int64))
    assert np.array_equal(col, Ho.indices.astype(np.int32))
    scale = max(np.max(np.abs(Ho.data)), 1e-300)
    assert np.max(np.abs(val - Ho.data)) <= 1e-14 * scale


@pytest.mark.parametrize("lvals,obc,pure,g,m,kat", CASES)
def test_sector_csr_energy(gpu, lvals, obc, pure, g, m, kat):
    print(f"case lvals={lvals} obc={obc} pure={pure}", flush=True)
    orc = QEDOracle(1, pure, lvals, obc)
    mod = _model(lvals, obc, pure)
    assert mod.sector_dim == orc.N
    assert np.array_equal(mod.sector_configs, orc.configs)
    assert np.array_equal(mod.loc_dims, orc.loc_dims)
    Hg = mod.build_Hamiltonian(g, m)
    Ho = orc.hamiltonian(g, m)
    assert Hg.is_complex == (not pure)
    _csr_match(Hg, Ho)
    k = min(4, orc.N - 1)
    ref = lowest_eigs(Ho, k)
    mod.diagonalize_Hamiltonian(k, "sparse")
    got = mod.H.Nenergies * np.prod(lvals)
    assert np.max(np.abs(got - ref)) < 1e-10, (got, ref)
    if kat is not None:
        assert abs(mod.H.Nenergies[0] - kat) < 1e-10
    assert np.all(mod.H.residuals < 1e-9 * np.maximum(1, np.abs(got)))


def test_c2_sector_csr_and_four_eigenvalues(gpu):
    """2D QED 4x2 PBC (N=140,647): bit-exact basis and CSR pattern, lowest 4
    eigenvalues at m=0.5 within 1e-10 for the literal couplings g in
    {0.5, 1, 2} and for g = {0.25, 4}, the squares the configuration's
    formula (g^2/2 E^2, 1/(2 g^2) plaquettes) maps them to; KAT at g=m=1."""
    orc = QEDOracle(1, False, [4, 2], [False, False])
    mod = _model([4, 2], [False, False], False)
    assert mod.sector_dim == 140647
    assert np.array_equal(mod.sector_configs, orc.configs)
    for g in (0.5, 1.0, 2.0, 0.25, 4.0):
        print(f"QED 4x2 g={g}", flush=True)
        Hg = mod.build_Hamiltonian(g, 0.5)
        Ho = orc.hamiltonian(g, 0.5)
        _csr_match(Hg, Ho)
        ref = lowest_eigs(Ho, 4)
        mod.diagonalize_Hamiltonian(4, "sparse")
        got = mod.H.Nenergies * 8
        assert np.max(np.abs(got - ref)) < 1e-10, (g, got, ref)
    mod.build_Hamiltonian(1.0, 1.0)
    mod.diagonalize_Hamiltonian(1, "sparse")
    assert abs(mod.H.Nenergies[0] - (-0.8830087027925408)) < 1e-10


def test_spmv_matches_scipy(gpu):
    import torch
    mod = _model([4, 2], [True, True], False)
    Hg = mod.build_Hamiltonian(1.0, 1.0)
    Hs = Hg.to_scipy()
    rng = np.random.default_rng(1)
    x = rng.standard_normal(Hg.n) + 1j * rng.standard_normal(Hg.n)
    y = Hg.matvec(torch.as_tensor(x, device="cuda")).cpu().numpy()
    ref = Hs @ x
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    X = rng.standard_normal((3, Hg.n)) + 1j * rng.standard_normal((3, Hg.n))
    Y = Hg.matvec(torch.as_tensor(X, device="cuda")).cpu().numpy()
    assert np.max(np.abs(Y - (Hs @ X.T).T)) <= 1e-13 * np.max(np.abs(ref))


def test_mixed_precision_lanczos(gpu):
    """FP32 SpMV with FP64 recurrence and FP64 refinement reaches |dE| <= 1e-10."""
    from edlgt_amd import PREC_MIXED
    orc = QEDOracle(1, False, [4, 2], [False, False])
    mod = _model([4, 2], [False, False], False)
    mod.build_Hamiltonian(1.0, 0.5)
    ref = lowest_eigs(orc.hamiltonian(1.0, 0.5), 2)
    mod.diagonalize_Hamiltonian(2, "sparse", precision=PREC_MIXED)
    got = mod.H.Nenergies * 8
    assert np.max(np.abs(got - ref)) < 1e-10
    assert mod.H.matvecs[1] > 0  # the FP32 phase ran


@pytest.mark.parametrize("lvals,obc,pure,g,m", [
    ([4, 2], [False, False], False, 1.0, 0.5),   # complex128 vectors
    ([3, 3], [False, False], True, 1.0, None),   # float64 vectors
    ([10], [True], False, 1.0, 1.0),
])
def test_mixed_precision_newton_one_eigenpair(gpu, lvals, obc, pure, g, m):
    """One eigenpair in mixed precision: loose FP32 Lanczos, then inexact
    Newton steps (FP64 residual, FP32 CG correction).  The lowest eigenvalue
    within 1e-10 of the oracle, the FP64 residual within tol, and the
    refinement path taken (a handful of FP64 products, not an FP64 Lanczos
    phase)."""
    from edlgt_amd import PREC_MIXED
    orc = QEDOracle(1, pure, lvals, obc)
    mod = _model(lvals, obc, pure)
    mod.build_Hamiltonian(g, m)
    ref = lowest_eigs(orc.hamiltonian(g, m), 1)
    mod.diagonalize_Hamiltonian(1, "sparse", precision=PREC_MIXED)
    e = float(mod.H.Nenergies[0] * np.prod(lvals))
    assert abs(e - ref[0]) < 1e-10, (e, ref[0])
    assert mod.H.residuals[0] <= 1e-9 * max(1.0, abs(e))
    mv64, mv32 = (int(v) for v in mod.H.matvecs)
    assert mv32 > 0 and mv64 <= 8, (mv64, mv32)
    # the returned vector is a unit vector
    assert abs(mod.H.Npsi[0].get_norm() - 1.0) < 1e-12


def test_lookup_and_initial_state(gpu):
    mod = _model([10], [True], False)
    cfg = mod.sector_configs
    idx = mod.basis.lookup(cfg[[0, 5, 196]])
    assert list(idx) == [0, 5, 196]
    bad = cfg[0].copy()
    bad[3] = (bad[3] + 1) % 5
    with pytest.raises(ValueError):
        mod.basis.lookup(bad)
    psi = mod.get_qmb_state_from_configs(cfg[[1, 2]])
    assert abs(float(psi.abs().pow(2).sum()) - 1.0) < 1e-14


def test_qed_3d_theta_sector_csr_energy(gpu):
    """3D pure PBC with the theta term (complex H): sector and CSR against the
    oracle, energy against the reference's KAT."""
    from edlgt_amd import QED_Model
    print("QED 2x2x2 PBC pure theta", flush=True)
    orc = QEDOracle(1, True, [2, 2, 2], [False, False, False])
    mod = QED_Model(spin=1, pure_theory=True, lvals=[2, 2, 2], has_obc=[False, False, False])
    assert mod.sector_dim == orc.N
    assert np.array_equal(mod.sector_configs, orc.configs)
    Hg = mod.build_Hamiltonian(2.3, theta=0.41)
    Ho = orc.hamiltonian(2.3, theta=0.41)
    _csr_match(Hg, Ho)
    mod.diagonalize_Hamiltonian(1, "sparse", tol=1e-12)
    assert abs(mod.H.Nenergies[0] - (-0.37555632787741716)) < 1e-10


@pytest.mark.parametrize("k", [270, 300, "full"])
def test_many_eigenpairs_above_512_states(gpu, k):
    """QED 2x2 PBC with matter (547 states, above the 512-state full-basis
    limit): n_eigs near or above half the sector, where a basis capped at
    N/2 vectors could not hold the wanted pairs, and n_eigs='full' -- every
    eigenvalue against the oracle's dense spectrum (eigsh accepts such k)."""
    orc = QEDOracle(1, False, [2, 2], [False, False])
    mod = _model([2, 2], [False, False], False)
    assert orc.N == 547
    mod.build_Hamiltonian(0.7, 0.3)
    w = np.linalg.eigvalsh(orc.hamiltonian(0.7, 0.3).toarray())
    mod.diagonalize_Hamiltonian(k, "sparse")
    got = mod.H.Nenergies * 4
    kk = orc.N if k == "full" else k
    assert got.size == kk
    assert np.max(np.abs(got - w[:kk])) < 1e-10
