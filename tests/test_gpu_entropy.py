"""GPU: entanglement entropy and reduced density matrices on the device
(edlgt_schmidt_gram + edlgt_hermitian_eigvals behind
QMB_state.entanglement_entropy / reduced_density_matrix, qmb_state.py:160-291).

Bars: squared Schmidt values within 1e-13 of the oracle's dense SVD (or its
block-by-block SVD where the dense matrix is large), entropy within 1e-12,
reduced density matrix within 1e-14 of the dense M M^H.  The SU(2) KAT
entropy (0.802957045143 at 1e-10) goes through the same path in
test_gpu_su2.py and test_gpu_workflows.py.
"""
import numpy as np
import pytest

from oracle.entropy import (block_schmidt_weights, dense_rdm, dense_schmidt_weights, entropy,
                            partition_maps)
from oracle.model import QEDOracle, SU2Oracle

pytestmark = pytest.mark.gpu


def _spectra_match(got, ref, tol=1e-13):
    got, ref = np.sort(got)[::-1], np.sort(ref)[::-1]
    n = min(got.size, ref.size)
    assert np.max(np.abs(got[:n] - ref[:n])) < tol
    assert np.all(np.abs(got[n:]) < tol) and np.all(np.abs(ref[n:]) < tol)


@pytest.mark.parametrize("real", [False, True])
def test_random_state_spectrum_entropy_rdm(gpu, real):
    """SU(2) 10 sites OBC, random complex (or real) state, four bipartitions."""
    import torch
    from edlgt_amd.models import QMB_state, schmidt_spectrum
    orc = SU2Oracle(False, [10], [True], sectors=[10])
    rng = np.random.default_rng(11)
    psi = rng.standard_normal(orc.N) + (0 if real else 1j * rng.standard_normal(orc.N))
    psi /= np.linalg.norm(psi)
    dev = torch.as_tensor(psi.real if real else psi, device="cuda")
    assert dev.dtype == (torch.float64 if real else torch.complex128)
    st = QMB_state(dev, [10], None)
    for keep in ([0, 1, 2, 3, 4], [0, 2, 5], [9], [3, 4, 5, 6]):
        print(f"SU(2) L=10 keep={keep} real={real}", flush=True)
        part = partition_maps(orc.configs, keep)
        pd = {tuple(sorted(keep)): part}
        ref = dense_schmidt_weights(psi, part)
        _spectra_match(schmidt_spectrum(dev, part), ref)
        assert abs(st.entanglement_entropy(keep, pd) - entropy(ref)) < 1e-12
        rho = st.reduced_density_matrix(keep, pd)
        assert np.max(np.abs(rho - dense_rdm(psi, part))) < 1e-14


def test_large_blocks_against_block_oracle(gpu):
    """SU(2) 12 sites, half chain: 13 blocks up to 141 x 141 (multi-tile Gram,
    141-step Jacobi sweeps); QED 4x2 PBC half lattice: 4,531 blocks."""
    import torch
    from edlgt_amd.models import schmidt_spectrum
    rng = np.random.default_rng(3)
    for name, orc, keep in (("SU(2) L=12", SU2Oracle(False, [12], [True], sectors=[12]), [0, 1, 2, 3, 4, 5]),
                            ("QED 4x2 PBC", QEDOracle(1, False, [4, 2], [False, False]), [0, 1, 2, 3])):
        print(f"{name} keep={keep}", flush=True)
        psi = rng.standard_normal(len(orc.configs)) + 1j * rng.standard_normal(len(orc.configs))
        psi /= np.linalg.norm(psi)
        part = partition_maps(orc.configs, keep)
        ref = block_schmidt_weights(psi, part)
        got = schmidt_spectrum(torch.as_tensor(psi, device="cuda"), part)
        _spectra_match(got, ref)
        assert abs(entropy(got) - entropy(ref)) < 1e-12
        assert abs(np.sum(got) - 1.0) < 1e-12


def test_ground_states_through_the_model(gpu):
    """Model ground states: QED 4x2 OBC (complex128) and pure QED 3x3 OBC
    (float64 vector), entropy of the engine's state against the oracle's
    dense SVD of the same vector."""
    from edlgt_amd import QED_Model
    for lvals, pure, m, keep in (([4, 2], False, 1.0, [0, 1, 2, 3]), ([3, 3], True, None, [0, 1, 2])):
        print(f"QED {lvals} pure={pure} keep={keep}", flush=True)
        mod = QED_Model(spin=1, pure_theory=pure, lvals=lvals, has_obc=[True, True])
        mod.build_Hamiltonian(1.0, m)
        mod.diagonalize_Hamiltonian(1, "sparse")
        st = mod.H.Npsi[0]
        mod._get_partition(keep)
        S = st.entanglement_entropy(keep, mod._partition_cache)
        part = partition_maps(mod.sector_configs, keep)
        assert S == pytest.approx(entropy(dense_schmidt_weights(st.psi, part)), abs=1e-12)
        assert S > 0.0


def test_c_abi_argument_checks(gpu):
    """A position outside the blocks and a short Gram buffer are errors (the
    scatter kernel checks every position; nothing is written out of bounds)."""
    import torch
    from edlgt_amd import _lib as L
    lib = L.load()
    psi = torch.ones(4, dtype=torch.complex128, device="cuda")
    blocks = np.array([[2, 2, 0, 0]], dtype=np.int64)
    G = torch.empty(4, dtype=torch.complex128, device="cuda")
    good = torch.tensor([0, 1, 2, 3], dtype=torch.int64, device="cuda")
    bad = torch.tensor([0, 1, 2, 4], dtype=torch.int64, device="cuda")
    L.check(lib.edlgt_schmidt_gram(L.ptr(psi), L.VALUE_C128, 4, L.ptr(good), 1, L.ptr(blocks), L.ptr(G), 4,
                                   L.stream_handle()))
    assert np.allclose(G.cpu().numpy(), [2, 2, 2, 2])
    with pytest.raises(ValueError):
        L.check(lib.edlgt_schmidt_gram(L.ptr(psi), L.VALUE_C128, 4, L.ptr(bad), 1, L.ptr(blocks), L.ptr(G), 4,
                                       L.stream_handle()))
    with pytest.raises(ValueError):
        L.check(lib.edlgt_schmidt_gram(L.ptr(psi), L.VALUE_C128, 4, L.ptr(good), 1, L.ptr(blocks), L.ptr(G), 3,
                                       L.stream_handle()))
    ev = np.zeros(2)
    L.check(lib.edlgt_schmidt_gram(L.ptr(psi), L.VALUE_C128, 4, L.ptr(good), 1, L.ptr(blocks), L.ptr(G), 4,
                                   L.stream_handle()))
    L.check(lib.edlgt_hermitian_eigvals(L.ptr(G), 4, 1, L.ptr(blocks), L.ptr(ev), 2, L.stream_handle()))
    assert np.allclose(np.sort(ev), [0.0, 4.0], atol=1e-14)
    with pytest.raises(ValueError):
        L.check(lib.edlgt_hermitian_eigvals(L.ptr(G), 4, 1, L.ptr(blocks), L.ptr(ev), 1, L.stream_handle()))
