"""GPU: QED_Model constructor options of the dressed-site basis.

link_symmetries=False (QED_model.py:131-141, 381-393): the sector keeps no
link constraints and H carries the penalty 100 E2 - 100 E_p E_m per link;
sector and CSR bit-exact against the oracle, the lowest levels equal to the
link-symmetric sector's (and the reference's KAT energy), the multi-term
(coupling sweep) form likewise.  get_only_bulk=True (QED_operators.py:410-538,
lattice_geometry.py:345-373): with periodic boundaries the same model, with
open ones an error (no border bases).  plaq_basis: not built (an error)."""
import numpy as np
import pytest

from oracle.model import QEDOracle, lowest_eigs

pytestmark = pytest.mark.gpu

CASES = [
    # lvals, has_obc, pure, g, m, KAT energy per site
    ([4], [True], False, 1.0, 1.0, -0.5691411466075982),
    ([4], [False], False, 1.0, 1.0, -0.5903881075118848),
    ([2, 2], [True, True], True, 1.0, None, None),
    ([2, 2], [False, False], True, 0.7, None, None),
    ([3, 2], [True, True], False, 0.8, 0.6, None),
]


def _csr_match(Hg, Ho):
    Ho = Ho.tocsr()
    rp, col, val = Hg.to_host()
    assert np.array_equal(rp, Ho.indptr.astype(np.int64))
    assert np.array_equal(col, Ho.indices.astype(np.int32))
    assert np.max(np.abs(val - Ho.data)) <= 1e-14 * max(np.max(np.abs(Ho.data)), 1e-300)


@pytest.mark.parametrize("lvals,obc,pure,g,m,kat", CASES)
def test_without_link_symmetries(gpu, lvals, obc, pure, g, m, kat):
    from edlgt_amd import QED_Model
    print(f"link_symmetries=False lvals={lvals} obc={obc} pure={pure}", flush=True)
    orc = QEDOracle(1, pure, lvals, obc, link_symmetries=False)
    ref = QEDOracle(1, pure, lvals, obc)
    mod = QED_Model(spin=1, pure_theory=pure, lvals=lvals, has_obc=obc, link_symmetries=False)
    assert mod.link_symmetries is False
    assert mod.sector_dim == orc.N > ref.N
    assert np.array_equal(mod.sector_configs, orc.configs)
    Hg = mod.build_Hamiltonian(g, m)
    _csr_match(Hg, orc.hamiltonian(g, m))
    k = min(2, ref.N - 1)
    mod.diagonalize_Hamiltonian(k, "sparse")
    e = mod.H.Nenergies * np.prod(lvals)
    assert np.max(np.abs(e - lowest_eigs(ref.hamiltonian(g, m), k))) < 1e-10
    if kat is not None:
        assert abs(mod.H.Nenergies[0] - kat) < 1e-10


def test_without_link_symmetries_coupling_sweep(gpu):
    from edlgt_amd import QED_Model
    orc = QEDOracle(1, False, [4], [True], link_symmetries=False)
    mod = QED_Model(spin=1, pure_theory=False, lvals=[4], has_obc=[True], link_symmetries=False)
    H3 = mod.build_multi_term(0.6)
    for g in (0.5, 1.0, 2.0):
        Ho = orc.hamiltonian(g, 0.6)
        d = abs(H3.to_scipy(coef=mod.multi_term_coefficients(g)) - Ho)
        assert (d.max() if d.nnz else 0.0) <= 1e-13 * max(1.0, abs(Ho).max())


def test_get_only_bulk_and_plaquette_basis(gpu):
    from edlgt_amd import QED_Model
    a = QED_Model(spin=1, pure_theory=False, lvals=[2, 2], has_obc=[False, False])
    b = QED_Model(spin=1, pure_theory=False, lvals=[2, 2], has_obc=[False, False], get_only_bulk=True)
    assert np.array_equal(a.sector_configs, b.sector_configs)
    a.build_Hamiltonian(0.7, 0.3)
    b.build_Hamiltonian(0.7, 0.3)
    a.diagonalize_Hamiltonian(2, "sparse")
    b.diagonalize_Hamiltonian(2, "sparse")
    assert np.max(np.abs(a.H.Nenergies - b.H.Nenergies)) < 1e-12
    with pytest.raises(ValueError):
        QED_Model(spin=1, pure_theory=False, lvals=[4], has_obc=[True], get_only_bulk=True)
    with pytest.raises(NotImplementedError):
        QED_Model(spin=1, pure_theory=True, lvals=[2, 2], has_obc=[False, False], plaq_basis=True)


def test_get_abelian_symmetry_sector_rebuilds_the_model_sectors(gpu):
    """QuantumModel.get_abelian_symmetry_sector (quantum_model.py:504-604) with
    the generators the reference's QED constructor passes (QED_model.py:118-199):
    the global N_tot = L/2 with the link constraints (matter), the link
    constraints with the origin-line flux constraints (pure, periodic 2D), and
    a global-only sector, against the constructor's sector and the oracle."""
    from edlgt_amd import QED_Model, SU2_Model
    from oracle.model import QEDOracle
    m = QED_Model(spin=1, pure_theory=False, lvals=[4, 2], has_obc=[False, False])
    want = m.sector_configs.copy()
    m.get_abelian_symmetry_sector(global_ops=[m.ops["N"]], global_sectors=[4],
                                  link_ops=[[m.ops[f"E_p{d}"], -m.ops[f"E_m{d}"]] for d in "xy"],
                                  link_sectors=[0, 0])
    assert np.array_equal(m.sector_configs, want)
    p = QED_Model(spin=1, pure_theory=True, lvals=[2, 2], has_obc=[False, False])
    want = p.sector_configs.copy()
    p.get_abelian_symmetry_sector(link_ops=[[p.ops[f"E_p{d}"], -p.ops[f"E_m{d}"]] for d in "xy"],
                                  link_sectors=[0, 0], nbody_ops=[p.ops["E_px"], p.ops["E_py"]],
                                  nbody_sectors=[0, 0],
                                  nbody_sites_list=[p.lattice.origin_line("y"), p.lattice.origin_line("x")],
                                  nbody_sym_type="U")
    assert np.array_equal(p.sector_configs, want)
    assert np.array_equal(want, QEDOracle(1, True, [2, 2], [False, False]).configs)
    s = SU2_Model(spin=0.5, pure_theory=False, sectors=[4], lvals=[4], has_obc=[True])
    s.get_abelian_symmetry_sector(global_ops=[s.ops["N_tot"]], global_sectors=[4])
    cfg = s.sector_configs
    nt = np.real(np.einsum("sii->si", s.ops["N_tot"]))
    assert np.all(np.rint(nt[np.arange(4)[None, :], cfg].sum(axis=1)) == 4)
    full = np.array(np.meshgrid(*[np.arange(d) for d in s.loc_dims], indexing="ij")).reshape(4, -1).T
    assert cfg.shape[0] == int(np.sum(np.rint(nt[np.arange(4)[None, :], full].sum(axis=1)) == 4))
