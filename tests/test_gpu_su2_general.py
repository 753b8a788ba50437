"""GPU: the generalised SU(2) model (any rishon truncation spin,
SU2_Model(use_generic_model=True) or spin >= 1: SU2_model.py:31-171, 357-484,
SU2_operators.py:302-592) and 3D SU(2) (hard-coded spin-1/2 set,
SU2_operators.py:35-300).

Bars: sector and CSR pattern bit-exact against the oracle, values within
1e-14 relative, lowest eigenvalues within 1e-10; the reference's own
consistency check (validation/su2/test_su2_hamiltonians.py: generalised and
hard-coded sets at spin 1/2 give the same sector and the same lowest levels
at g = m = 0.1, atol 1e-10) on the engine, 1D to 3D."""
import numpy as np
import pytest

from oracle.model import SU2Oracle, lowest_eigs

pytestmark = pytest.mark.gpu


def _csr_match(Hg, Ho):
    Ho = Ho.tocsr()
    rp, col, val = Hg.to_host()
    assert np.array_equal(rp, Ho.indptr.astype(np.int64))
    assert np.array_equal(col, Ho.indices.astype(np.int32))
    assert np.max(np.abs(val - Ho.data)) <= 1e-14 * max(np.max(np.abs(Ho.data)), 1e-300)


CASES = [
    # lvals, has_obc, pure, spin, generic
    ([4], [True], False, 1.0, True),
    ([4], [False], False, 1.0, True),
    ([6], [True], False, 1.0, True),
    ([4], [True], False, 1.5, True),
    ([2, 2], [True, True], False, 1.0, True),
    ([2, 2], [False, False], False, 1.0, True),  # 95,252 states, up to 356 raw entries per row
    ([2, 2], [False, False], True, 1.0, True),
    ([3, 2], [True, True], True, 1.0, True),
    ([2, 2], [False, False], False, 0.5, True),
    ([2, 2, 2], [True] * 3, True, 0.5, False),
    ([2, 2, 2], [True] * 3, False, 0.5, False),
]


@pytest.mark.parametrize("lvals,obc,pure,spin,generic", CASES)
def test_sector_csr_energy(gpu, lvals, obc, pure, spin, generic):
    from edlgt_amd import SU2_Model
    print(f"SU(2) lvals={lvals} obc={obc} pure={pure} spin={spin} generic={generic}", flush=True)
    n = int(np.prod(lvals))
    sec = None if pure else [n]
    m = None if pure else 0.7
    orc = SU2Oracle(pure, lvals, obc, sectors=sec, spin=spin, generic=generic)
    mod = SU2_Model(spin=spin, pure_theory=pure, sectors=sec, lvals=lvals, has_obc=obc, use_generic_model=generic)
    assert mod.use_generic_model == (generic or spin >= 1)
    assert np.array_equal(mod.loc_dims, orc.loc_dims)
    assert mod.sector_dim == orc.N
    assert np.array_equal(mod.sector_configs, orc.configs)
    Hg = mod.build_Hamiltonian(1.3, m)
    Ho = orc.hamiltonian(1.3, m)
    _csr_match(Hg, Ho)
    k = min(3, orc.N - 1)
    mod.diagonalize_Hamiltonian(k, "sparse")
    ref = lowest_eigs(Ho, k) if orc.N > 600 else np.linalg.eigvalsh(Ho.toarray())[:k]
    assert np.max(np.abs(mod.H.Nenergies * n - ref)) < 1e-10


@pytest.mark.parametrize("lvals,obc,pure", [([4], [True], False), ([4], [False], False),
                                            ([2, 2], [True, True], False), ([2, 2], [False, False], False),
                                            ([2, 2], [True, True], True), ([2, 2], [False, False], True),
                                            ([2, 2, 2], [True] * 3, False), ([2, 2, 2], [True] * 3, True)])
def test_generalised_equals_hard_coded(gpu, lvals, obc, pure):
    """compare_SU2_models (workflows/su2.py:36-89) on the engine."""
    from edlgt_amd import SU2_Model
    print(f"compare SU(2) models lvals={lvals} obc={obc} pure={pure}", flush=True)
    n = int(np.prod(lvals))
    sec = None if pure else [n]
    m = None if pure else 0.1
    base = SU2_Model(spin=0.5, pure_theory=pure, sectors=sec, lvals=lvals, has_obc=obc)
    gen = SU2_Model(spin=0.5, pure_theory=pure, sectors=sec, lvals=lvals, has_obc=obc, use_generic_model=True)
    assert np.array_equal(base.sector_configs, gen.sector_configs)
    k = min(4 if len(lvals) <= 2 else 2, base.sector_dim - 1)
    if k < 1:
        return
    base.build_Hamiltonian(0.1, m)
    gen.build_Hamiltonian(0.1, m)
    base.diagonalize_Hamiltonian(k, "sparse")
    gen.diagonalize_Hamiltonian(k, "sparse")
    assert np.max(np.abs(base.H.Nenergies - gen.H.Nenergies)) < 1e-10


def test_arguments(gpu):
    from edlgt_amd import SU2_Model
    with pytest.raises(ValueError):
        SU2_Model(spin=0.7, pure_theory=True, lvals=[2, 2], has_obc=[True, True])
    with pytest.raises(NotImplementedError):
        SU2_Model(spin=1.0, pure_theory=True, lvals=[2, 2, 2], has_obc=[True] * 3)
    with pytest.raises(NotImplementedError):
        SU2_Model(spin=1.5, pure_theory=False, sectors=[4], lvals=[2, 2], has_obc=[True, True])  # 600 states
    gen = SU2_Model(spin=1.0, pure_theory=False, sectors=[4], lvals=[4], has_obc=[True])
    with pytest.raises(NotImplementedError):
        gen.overlap_QMB_state("V")


@pytest.mark.parametrize("pure", [True, False])
def test_3d_theta_term(gpu, pure):
    """3D SU(2) with the topological term (SU2_model.py:303-336): sector and
    CSR bit-exact against the oracle, complex H, energies within 1e-10."""
    from edlgt_amd import SU2_Model
    print(f"SU(2) 3D theta pure={pure}", flush=True)
    sec = None if pure else [8]
    m = None if pure else 0.5
    orc = SU2Oracle(pure, [2, 2, 2], [True] * 3, sectors=sec)
    mod = SU2_Model(spin=0.5, pure_theory=pure, sectors=sec, lvals=[2, 2, 2], has_obc=[True] * 3)
    Hg = mod.build_Hamiltonian(1.1, m, theta=0.45)
    assert Hg.is_complex
    Ho = orc.hamiltonian(1.1, m, theta=0.45)
    _csr_match(Hg, Ho)
    k = min(3, orc.N - 1)
    mod.diagonalize_Hamiltonian(k, "sparse")
    ref = lowest_eigs(Ho, k) if orc.N > 600 else np.linalg.eigvalsh(Ho.toarray())[:k]
    assert np.max(np.abs(mod.H.Nenergies * 8 - ref)) < 1e-10


BG_CASES = [
    # lvals, has_obc, pure, bg_list, noise, spin, generic
    ([4], [True], False, [0.5, 0, 0, 0.5], 0.0, 0.5, False),
    ([4], [False], False, [0.5, 0, 0.5, 0], 0.0, 0.5, False),
    ([4], [False], False, [0.5, 0, 0.5, 0], 0.3, 0.5, False),   # Gauss-law-violating hopping
    ([2, 2], [True, True], True, [0.5, 0, 0, 0.5], 0.0, 0.5, False),
    ([2, 2], [True, True], False, [0.5, 0, 0, 0.5], 0.0, 0.5, False),
    # the generalised set with background slots (SU2_operators.py:580-590)
    ([4], [False], False, [0.5, 0, 0.5, 0], 0.0, 0.5, True),
    ([4], [True], False, [0.5, 0, 1, 0.5], 0.0, 1.0, True),
    ([4], [False], False, [1, 0, 1, 0], 0.0, 1.0, True),
    ([2, 2], [False, False], True, [0, 1, 0, 1], 0.0, 1.0, True),
]


@pytest.mark.parametrize("lvals,obc,pure,bg,noise,spin,generic", BG_CASES)
def test_background_charges(gpu, lvals, obc, pure, bg, noise, spin, generic):
    """SU2_Model(bg_list=...) (SU2_model.py:63-104, 338-354;
    quantum_model.py:440-466): sector and CSR bit-exact against the oracle,
    energies within 1e-10."""
    from edlgt_amd import SU2_Model
    print(f"SU(2) background lvals={lvals} obc={obc} pure={pure} bg={bg} noise={noise} spin={spin} "
          f"generic={generic}", flush=True)
    n = int(np.prod(lvals))
    sec = None if pure else [n]
    m = None if pure else 0.8
    orc = SU2Oracle(pure, lvals, obc, sectors=sec, spin=spin, generic=generic, bg_list=bg)
    mod = SU2_Model(spin=spin, pure_theory=pure, sectors=sec, lvals=lvals, has_obc=obc, bg_list=bg,
                    use_generic_model=generic)
    assert mod.background == max(bg) and mod.bg_list == [float(b) for b in bg]
    assert mod.use_generic_model == (generic or spin >= 1)
    assert np.array_equal(mod.loc_dims, orc.loc_dims)
    assert np.array_equal(mod.sector_configs, orc.configs)
    Hg = mod.build_Hamiltonian(1.1, m, lambda_noise=noise)
    Ho = orc.hamiltonian(1.1, m, noise=noise)
    _csr_match(Hg, Ho)
    k = min(2, orc.N - 1) if orc.N > 1 else 1
    mod.diagonalize_Hamiltonian(k, "sparse")
    ref = lowest_eigs(Ho, k) if orc.N > 600 else np.linalg.eigvalsh(Ho.toarray())[:k]
    assert np.max(np.abs(mod.H.Nenergies * n - ref)) < 1e-10


def test_background_spin1_flux_string_closed_form(gpu):
    """Pure chain at spin-1 truncation, charges 1/2, 1, 1/2 at sites 0, 2, 3
    (OBC): every link carries j = 1/2, one state of energy
    (8g/3) * 3 * (3/4)."""
    from edlgt_amd import SU2_Model
    mod = SU2_Model(spin=1, pure_theory=True, lvals=[4], has_obc=[True], bg_list=[0.5, 0, 1, 0.5])
    assert mod.sector_dim == 1 and mod.use_generic_model
    mod.build_Hamiltonian(0.7)
    mod.diagonalize_Hamiltonian(1, "sparse")
    assert abs(mod.H.Nenergies[0] * 4 - (8 * 0.7 / 3) * 3 * 0.75) < 1e-12


def test_background_flux_string_closed_form(gpu):
    """Pure chain, charges 1/2 at sites 1 and 4 of 6 (OBC): one state, energy
    2 g per link between the charges."""
    from edlgt_amd import SU2_Model
    mod = SU2_Model(spin=0.5, pure_theory=True, lvals=[6], has_obc=[True], bg_list=[0, 0.5, 0, 0, 0.5, 0])
    assert mod.sector_dim == 1
    mod.build_Hamiltonian(1.3)
    mod.diagonalize_Hamiltonian(1, "sparse")
    assert abs(mod.H.Nenergies[0] * 6 - 2 * 1.3 * 3) < 1e-12
    with pytest.raises(ValueError):
        SU2_Model(spin=0.5, pure_theory=True, lvals=[4], has_obc=[True], bg_list=[0.3, 0, 0, 0])


def test_named_builders(gpu):
    """The builders by their reference names (SU2_model.py:200, 357;
    QED_model.py:235, 396) give the dispatched matrix, and refuse the other
    operator set."""
    from edlgt_amd import QED_Model, SU2_Model
    gen = SU2_Model(spin=1.0, pure_theory=False, sectors=[4], lvals=[4], has_obc=[True])
    a = gen.build_Hamiltonian(1.3, 0.7).to_host()
    b = gen.build_gen_Hamiltonian(1.3, 0.7).to_host()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    with pytest.raises(ValueError):
        gen.build_base_Hamiltonian(1.3, 0.7)
    base = SU2_Model(spin=0.5, pure_theory=False, sectors=[4], lvals=[4], has_obc=[True])
    a = base.build_Hamiltonian(1.3, 0.7).to_host()
    b = base.build_base_Hamiltonian(1.3, 0.7).to_host()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    with pytest.raises(ValueError):
        base.build_gen_Hamiltonian(1.3, 0.7)
    q = QED_Model(spin=1, pure_theory=False, lvals=[4], has_obc=[True])
    a = q.build_Hamiltonian(1.0, 1.0).to_host()
    b = q.build_truncated_Hamiltonian(1.0, 1.0).to_host()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    with pytest.raises(ValueError):
        q.build_integrated_Hamiltonian(1.0, 1.0)
