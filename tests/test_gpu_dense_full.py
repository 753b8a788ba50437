"""GPU: n_eigs="full" above the Lanczos full-basis limit (4096 states) takes
the dense route the reference takes (qmb_hamiltonian.py:293-301), on the
device: the 10-site open SU(2) chain with matter (N_tot = 10, 8,953 states,
complex H) factored by rocSOLVER.  Pinned without a host factorization of
the 8,953 x 8,953 matrix:
  * the lowest eigenvalues against the oracle's eigsh at 1e-10;
  * sum_i E_i = Tr H and sum_i E_i^2 = ||H||_F^2 (the oracle's matrix) at
    1e-11 relative;
  * residuals |H v_i - E_i v_i| through the engine's SpMV for 64 eigenvectors
    spread over the spectrum, and V V^H = I, at 1e-10;
  * the real-gauge solve gives the same spectrum and, mapped back, eigenvectors
    of the complex H;
  * the infinite-temperature closed forms of the thermal helpers over the
    dense spectrum (Z(0) = N, <e>_0 = Tr H / (N L)) and of the canonical average.
A matrix-free operator has no dense form (error)."""
import numpy as np
import pytest

from oracle.model import SU2Oracle, lowest_eigs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def chain(gpu):
    from edlgt_amd import SU2_Model
    orc = SU2Oracle(False, [10], [True], sectors=[10])
    Ho = orc.hamiltonian(1.0, 1.0)
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[10], lvals=[10], has_obc=[True])
    mod.build_Hamiltonian(1.0, 1.0)
    return orc, Ho, mod


def test_dense_full_spectrum(chain):
    import torch
    orc, Ho, mod = chain
    print(f"dense full solve, N = {orc.N}", flush=True)
    assert orc.N == 8953 and mod.sector_dim == orc.N
    mod.diagonalize_Hamiltonian("full", "sparse")
    E = mod.H.energies
    assert mod.H.n_eigs == orc.N and len(mod.H.Npsi) == orc.N and np.all(np.diff(E) >= 0)
    assert np.max(np.abs(E[:6] - lowest_eigs(Ho, 6))) < 1e-10
    tr = float(np.real(Ho.diagonal().sum()))
    fro2 = float(np.sum(np.abs(Ho.data) ** 2))
    assert abs(E.sum() - tr) < 1e-11 * max(1.0, np.abs(E).sum())
    assert abs(np.sum(E ** 2) - fro2) < 1e-11 * fro2
    V = torch.stack([mod.H.Npsi[i].psi_device for i in range(orc.N)])
    G = V @ V.conj().transpose(0, 1)
    G.diagonal().sub_(1.0)
    assert float(G.abs().max()) < 1e-10
    scale = float(np.max(np.abs(E)))
    for i in np.linspace(0, orc.N - 1, 64).astype(int):
        v = V[i].contiguous()
        r = mod.H.Ham.matvec(v) - E[i] * v
        assert float(torch.linalg.vector_norm(r)) < 1e-10 * scale, i
    # the thermal helpers over the stored per-site energies; canonical at beta = 0
    assert mod.H.partition_function(0.0) == pytest.approx(orc.N, rel=1e-13)
    assert abs(mod.H.thermal_average(0.0) - tr / (orc.N * 10)) < 1e-11
    diag = np.real(np.einsum("sii->si", orc.ops["N_single"]))
    inf_t = float(np.mean(diag[np.arange(10)[None, :], orc.configs].sum(axis=1))) / 10
    assert abs(mod.canonical_avg("N_single", 0.0) - inf_t) < 1e-11


def test_dense_full_real_gauge_and_errors(chain):
    import torch
    from edlgt_amd import SU2_Model
    orc, Ho, mod = chain
    print("dense full solve in the real gauge", flush=True)
    m2 = SU2_Model(spin=0.5, pure_theory=False, sectors=[10], lvals=[10], has_obc=[True])
    m2.build_Hamiltonian(1.0, 1.0)
    m2.diagonalize_Hamiltonian("full", "sparse", real_gauge=True)
    assert m2.H.real_gauge
    E = m2.H.energies
    assert np.max(np.abs(E[:6] - lowest_eigs(Ho, 6))) < 1e-10
    assert abs(E.sum() - float(np.real(Ho.diagonal().sum()))) < 1e-11 * np.abs(E).sum()
    scale = float(np.max(np.abs(E)))
    for i in (0, 1, 4476, orc.N - 1):
        v = m2.H.Npsi[i].psi_device.to(torch.complex128).contiguous()
        r = m2.H.Ham.matvec(v) - E[i] * v
        assert float(torch.linalg.vector_norm(r)) < 1e-10 * scale, i
    del m2
    mf = SU2_Model(spin=0.5, pure_theory=False, sectors=[10], lvals=[10], has_obc=[True])
    mf.matrix_free = True
    mf.build_Hamiltonian(1.0, 1.0)
    with pytest.raises(Exception, match="matrix-free"):
        mf.diagonalize_Hamiltonian("full", "sparse")
