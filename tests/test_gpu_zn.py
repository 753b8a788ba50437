"""GPU parity for the Z_n model: sector bit-exact, CSR pattern bit-exact and
values within 1e-14, eigenvalues within 1e-10 of the oracle; 128-bit keys on
the 4x4 lattice against the independent count 3^15 and the J=0 closed form."""
import numpy as np
import pytest

from oracle.model import ZnOracle, lowest_eigs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lv,obc,fs", [([2, 2], [False, False], None), ([3, 3], [False, False], (0, 0)),
                                       ([3, 3], [False, False], None), ([3, 2], [False, False], (1, 2)),
                                       ([4, 3], [False, False], (0, 0))])
def test_zn_parity(gpu, lv, obc, fs):
    from edlgt_amd import Zn_Model
    print(f"Zn lvals={lv} obc={obc} flux={fs}", flush=True)
    orc = ZnOracle(3, lv, obc, fs)
    mod = Zn_Model(3, lvals=lv, has_obc=obc, flux_sector=fs)
    assert mod.sector_dim == orc.N
    assert np.array_equal(mod.sector_configs, orc.configs)
    J, h = 1.0, 0.8
    Hg = mod.build_Hamiltonian(J, h)
    Ho = orc.hamiltonian(J, h)
    assert not Hg.is_complex
    rp, col, val = Hg.to_host()
    assert np.array_equal(rp, Ho.indptr) and np.array_equal(col, Ho.indices)
    assert np.max(np.abs(val - Ho.data)) <= 1e-14 * np.max(np.abs(Ho.data))
    k = min(4, orc.N - 1)
    mod.diagonalize_Hamiltonian(k, "sparse", tol=1e-12)
    ref = lowest_eigs(Ho, k)
    assert np.max(np.abs(mod.H.Nenergies * np.prod(lv) - ref)) < 1e-10


def test_zn_4x4_wide_keys(gpu):
    from edlgt_amd import Zn_Model
    mod = Zn_Model(3, lvals=[4, 4], has_obc=[False, False], flux_sector=(0, 0))
    assert mod.sector_dim == 3 ** 15
    keys = mod.basis.keys()
    assert len(keys) == 3 ** 15 and all(keys[i] < keys[i + 1] for i in range(0, 3 ** 15 - 1, 99991))
    cfg = mod.sector_configs
    idx = mod.basis.lookup(cfg[[0, 12345, 3 ** 15 - 1]])
    assert list(idx) == [0, 12345, 3 ** 15 - 1]
    mod.build_Hamiltonian(0.0, 1.0)
    mod.diagonalize_Hamiltonian(1, "sparse", tol=1e-12)
    assert abs(mod.H.Nenergies[0] * 16 - (-2 * 32)) < 1e-9
