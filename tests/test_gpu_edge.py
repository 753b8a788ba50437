"""GPU edge cases: an empty symmetry sector, a one-state sector, a sector
solved in full (k = N), invalid requests that must raise the reference's
exception types, observables/expectation values on tiny sectors, and the
device start vector (numpy's PCG64 stream)."""
import numpy as np
import pytest

from oracle.model import SU2Oracle

pytestmark = pytest.mark.gpu


def _su2(lv, obc, n):
    from edlgt_amd import SU2_Model
    return SU2_Model(spin=0.5, pure_theory=False, sectors=[n], lvals=lv, has_obc=obc)


def test_empty_sector(gpu):
    """N_tot = 100 on 4 sites has no states: SU2_Model raises ValueError, as
    the reference does (SU2_model.py:170-171)."""
    with pytest.raises(ValueError, match="No configurations"):
        _su2([4], [True], 100)


def test_one_state_sector(gpu):
    """N_tot = 0 on an open 2-site chain: one state, its diagonal is the energy."""
    mod = _su2([2], [True], 0)
    orc = SU2Oracle(False, [2], [True], sectors=[0])
    assert mod.sector_dim == 1 and np.array_equal(mod.sector_configs, orc.configs)
    mod.build_Hamiltonian(1.0, 1.0)
    mod.diagonalize_Hamiltonian(1, "sparse")
    e = orc.hamiltonian(1.0, 1.0).toarray()[0, 0].real
    assert abs(mod.H.energies[0] - e) < 1e-12
    mod.get_observables(["N_tot"])
    mod.measure_observables(0)
    assert np.allclose(mod.res["N_tot"], 0.0)


def test_full_spectrum_of_a_small_sector(gpu):
    """k = N on a 3-state sector: every eigenvalue, against dense eigvalsh."""
    mod = _su2([2], [True], 2)
    orc = SU2Oracle(False, [2], [True], sectors=[2])
    assert mod.sector_dim == 3
    mod.build_Hamiltonian(0.7, 1.3)
    mod.diagonalize_Hamiltonian(3, "sparse")
    ref = np.linalg.eigvalsh(orc.hamiltonian(0.7, 1.3).toarray())
    assert np.max(np.abs(mod.H.energies - ref)) < 1e-12
    with pytest.raises(ValueError):
        mod.diagonalize_Hamiltonian(4, "sparse")
    with pytest.raises(ValueError):
        mod.diagonalize_Hamiltonian(0, "sparse")
    # n_eigs="full" (the reference's dense path) gives every eigenpair
    mod.diagonalize_Hamiltonian("full", "sparse")
    assert np.max(np.abs(mod.H.energies - ref)) < 1e-12 and len(mod.H.Npsi) == 3
    big = _su2([6], [False], 6)
    big.build_Hamiltonian(1.0, 5.0)
    orc6 = SU2Oracle(False, [6], [False], sectors=[6])
    big.diagonalize_Hamiltonian("full", "sparse")
    ref6 = np.linalg.eigvalsh(orc6.hamiltonian(1.0, 5.0).toarray())
    assert np.max(np.abs(big.H.energies - ref6)) < 1e-10


def test_invalid_requests_raise(gpu):
    import torch
    mod = _su2([6], [False], 6)
    mod.build_Hamiltonian(1.0, 5.0)
    with pytest.raises(ValueError):
        mod.get_qmb_state_from_configs([[5] * 6])  # not in the sector
    psi0 = mod.get_qmb_state_from_configs([mod.overlap_QMB_state("V")])
    with pytest.raises(ValueError):
        mod.time_evolution_Hamiltonian(psi0, [0.0])  # one time point
    with pytest.raises(ValueError):
        mod.time_evolution_Hamiltonian(psi0, [0.0, 0.1, 0.3])  # not uniform
    with pytest.raises(ValueError):
        mod.time_evolution_Hamiltonian(torch.zeros(3, dtype=torch.complex128), [0.0, 0.1])
    with pytest.raises(ValueError):
        mod.get_observables(nbody_obs=[["Qpx_dag", "Qmx"]], nbody_dist=[[(1,), (2,)]])


def test_expectation_of_a_ket_outside_the_sector_raises(gpu):
    """A single hopping factor leaves the Gauss-law sector: an error, not a
    value (the reference's lookup would return -1 and read psi[-1])."""
    from edlgt_amd.device import expval_terms
    mod = _su2([6], [False], 6)
    mod.build_Hamiltonian(1.0, 5.0)
    mod.diagonalize_Hamiltonian(1, "sparse")
    psi = mod.H.Npsi[0].psi_device
    tab = np.stack([mod.ops["Qpx_dag"][0]])
    with pytest.raises(ValueError):
        expval_terms(mod.basis, tab, [dict(sites=[0, 2], tables=[0, 0])], psi)


@pytest.mark.parametrize("seed", [0, 7, 2**40 + 3])
def test_device_start_vector_is_numpy_default_rng(gpu, seed):
    """edlgt_random_uniform reproduces numpy.random.default_rng(seed).random(n)
    bit for bit (PCG64 jump-ahead per thread), for lengths around the
    4096-value block and a ragged large one; complex output in the real part."""
    import torch
    from edlgt_amd.device import random_uniform
    for n in (1, 255, 4096, 4097, 1_000_003):
        want = np.random.default_rng(seed).random(n)
        got = random_uniform(n, complex_values=False, seed=seed).cpu().numpy()
        assert np.array_equal(got, want), (seed, n)
    gc = random_uniform(9001, complex_values=True, seed=seed).cpu().numpy()
    assert np.array_equal(gc.real, np.random.default_rng(seed).random(9001)) and not gc.imag.any()
    assert random_uniform(0, seed=seed).numel() == 0
    torch.cuda.synchronize()


def test_default_start_equals_host_start(gpu):
    """diagonalize without v0 (device-generated default_rng(0) start) gives
    the energies of the same solve started from the host vector."""
    mod = _su2([6], [True], 6)
    mod.build_Hamiltonian(1.0, 1.0)
    mod.diagonalize_Hamiltonian(3, "sparse")
    e_dev = np.array(mod.H.energies)
    v0 = np.random.default_rng(0).random(mod.sector_dim)
    mod.H.diagonalize(3, "sparse", mod.loc_dims, v0=v0)
    assert np.allclose(e_dev, mod.H.energies, atol=1e-11, rtol=0)


def test_repeated_placements_are_summed(gpu):
    """Two raw entries of one row reaching the same ket (a placement listed
    twice) are summed into one stored entry: the assembly's table-predicted
    row counts are then too high, the fill pass detects it and the rows are
    counted by enumeration.  Checked against the dense sum of Kronecker
    products on the full 3-site product space (27 states)."""
    import scipy.sparse as sp
    from edlgt_amd.device import CSRMatrix, SectorBasis
    import edlgt_amd._lib as L
    rng = np.random.default_rng(5)
    A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    D = np.diag(rng.normal(size=3)).astype(complex)
    tables = np.stack([A, D])
    pl = [dict(sites=[0, 1], tables=[0, 0], strength=0.7),
          dict(sites=[0, 1], tables=[0, 0], strength=0.7),
          dict(sites=[1, 2], tables=[0, 0], strength=0.3, dagger=True),
          dict(sites=[2], tables=[1], strength=1.0)]
    basis = SectorBasis([3, 3, 3])
    H = CSRMatrix(basis, tables, pl, L.VALUE_C128)
    rp, col, val = H.to_host()
    got = sp.csr_matrix((val, col, rp), shape=(27, 27)).toarray()
    I = np.eye(3)
    want = 1.4 * np.kron(np.kron(A, A), I) + (0.3 * np.kron(I, np.kron(A, A))).conj().T \
        + np.kron(np.kron(I, I), D)
    assert np.allclose(got, want, atol=1e-13)
    # one stored entry per distinct position, columns ascending in every row
    assert H.nnz == np.count_nonzero(np.abs(want) > 1e-14)
    assert all(np.all(np.diff(col[rp[r]:rp[r + 1]]) > 0) for r in range(27))
    # the table-only count cannot see the collisions: an upper bound here
    from edlgt_amd.device import stored_entries
    assert stored_entries(basis, tables, pl) > H.nnz


def _kron_chain(factors):
    import scipy.sparse as sp
    out = sp.csr_matrix(np.ones((1, 1), complex))
    for f in factors:
        out = sp.kron(out, sp.csr_matrix(f), format="csr")
    return out


def test_rows_above_1024_raw_entries(gpu):
    """Rows whose raw entries exceed the first assembly pass's 1024 LDS slots
    are listed and assembled again with 4096 slots.  Six 4-state sites, 4-site
    placements of a table whose state 0 has four transitions and the others
    two: a row has 96 (no site in state 0) to 1536 (all in state 0) raw
    entries, so both passes run, with the exact-count fallback (overlapping
    placements of one term reach the same kets).  Checked against the sum of
    Kronecker products on the full 4096-state space."""
    from edlgt_amd.device import CSRMatrix, SectorBasis
    import edlgt_amd._lib as L
    rng = np.random.default_rng(11)
    A = np.zeros((4, 4), complex)
    A[0, :] = rng.normal(size=4) + 1j * rng.normal(size=4)
    A[:, 0] = rng.normal(size=4) + 1j * rng.normal(size=4)
    for a in range(1, 4):
        A[a, a] = rng.normal() + 1j * rng.normal()
    D = np.diag(rng.normal(size=4)).astype(complex)
    I = np.eye(4)
    pl, want = [], None
    for w, s in zip(range(3), (0.7, -0.4, 1.3)):
        pl.append(dict(sites=list(range(w, w + 4)), tables=[0] * 4, strength=s))
        pl.append(dict(sites=list(range(w, w + 4)), tables=[0] * 4, strength=0.5 * s, dagger=True))
        term = _kron_chain([I] * w + [A] * 4 + [I] * (2 - w))
        want = (s * term if want is None else want + s * term) + (0.5 * s * term).conj().T
    for site in range(6):
        pl.append(dict(sites=[site], tables=[1], strength=0.25))
        want = want + 0.25 * _kron_chain([I] * site + [D] + [I] * (5 - site))
    H = CSRMatrix(SectorBasis([4] * 6), np.stack([A, D]), pl, L.VALUE_C128)
    rp, col, val = H.to_host()
    want = want.tocsr()
    want.sum_duplicates()
    want.eliminate_zeros()
    want.sort_indices()
    assert np.array_equal(rp, want.indptr) and np.array_equal(col, want.indices)
    assert np.max(np.abs(val - want.data)) <= 1e-13 * np.max(np.abs(want.data))
    # raw entries per row: both directions of each window, product of the
    # factors' transition counts (4 from state 0, else 2): both passes ran
    per_site = np.where(np.indices([4] * 6).reshape(6, -1).T == 0, 4, 2)
    raw = sum(2 * per_site[:, w:w + 4].prod(axis=1) for w in range(3))
    assert raw.max() == 1536 and (raw > 1024).any() and (raw <= 1024).any()


def test_rows_above_4096_raw_entries_raise(gpu):
    """More than 4096 raw entries in one row is an error (the second pass's
    LDS bound), not a silent truncation: two 5-site placements of a dense
    4x4 table and their conjugates give 4096 per row, one 2-site placement
    more crosses the bound."""
    from edlgt_amd.device import CSRMatrix, SectorBasis
    import edlgt_amd._lib as L
    A = np.random.default_rng(12).normal(size=(4, 4)).astype(complex)
    pl = []
    for w in range(2):
        pl.append(dict(sites=list(range(w, w + 5)), tables=[0] * 5, strength=1.0))
        pl.append(dict(sites=list(range(w, w + 5)), tables=[0] * 5, strength=1.0, dagger=True))
    pl.append(dict(sites=[0, 1], tables=[0, 0], strength=1.0))
    with pytest.raises(OverflowError, match="4096 raw entries"):
        CSRMatrix(SectorBasis([4] * 6), A[None], pl, L.VALUE_C128)
