"""GPU parity of the stabilizer Renyi-2 entropy (QMB_state.
stabilizer_renyi_entropy, qmb_state.py:393-465) against the oracle's
restatement of the reference's algorithm (oracle/magic.py, pinned by closed
forms in tests/test_magic.py) on the same state vectors: the ground states of
the 6-site periodic SU(2) chain (tail tolerances 1e-2 and 1e-3: supports of
27 and 64 states) and of the 2x2 periodic QED lattice (about 240 states and
30,000 X-strings), SRE2 at 1e-10 with the same support, discarded weight and string
count; and the closed forms on the full 6-qubit product space (a product of
cos(t)|0> + sin(t)|1> factors has SRE2 = -sum log((a^2 + b^2)^2 + 12 a^2 b^2),
a basis state and the uniform superposition 0)."""
import itertools

import numpy as np
import pytest

from oracle import magic
from oracle.model import QEDOracle, SU2Oracle

pytestmark = pytest.mark.gpu


def _check(mod, orc, delta):
    psi = mod.H.Npsi[0].psi
    assert np.array_equal(mod.sector_configs, orc.configs)
    want = magic.sre2(psi, orc.loc_dims, orc.configs, delta)
    got = mod.basis.sre2(mod.H.Npsi[0].psi_device, delta)
    assert abs(got[0] - want[0]) < 1e-10, (got, want)
    assert got[1] == want[2] and got[3] == want[1]
    assert abs(got[2] - want[3]) < 1e-14
    return got


@pytest.mark.parametrize("delta", [1e-2, 1e-3])
def test_su2_ground_state(gpu, delta):
    from edlgt_amd import SU2_Model
    print(f"SRE2 SU(2) L=6 delta={delta}", flush=True)
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    mod.build_Hamiltonian(1.0, 5.0)
    mod.diagonalize_Hamiltonian(1, "sparse", tol=1e-12)
    got = _check(mod, orc, delta)
    # the reference's entry point: the state and the model's sector_configs
    assert mod.H.Npsi[0].stabilizer_renyi_entropy(mod.sector_configs, delta) == got[0]
    with pytest.raises(ValueError):
        mod.H.Npsi[0].stabilizer_renyi_entropy(np.array(mod.sector_configs), delta)
    with pytest.raises(ValueError):
        mod.basis.sre2(mod.H.Npsi[0].psi_device, 0.0)


def test_qed_ground_state(gpu):
    from edlgt_amd import QED_Model
    print("SRE2 QED 2x2", flush=True)
    orc = QEDOracle(1, False, [2, 2], [False, False])
    mod = QED_Model(spin=1, pure_theory=False, lvals=[2, 2], has_obc=[False, False])
    mod.build_Hamiltonian(1.0, 0.5)
    mod.diagonalize_Hamiltonian(1, "sparse", tol=1e-12)
    got = _check(mod, orc, 1e-2)
    # which of the equal-weight states at the cut are kept follows the
    # vector's last bits: only the scale of the case is fixed here
    assert got[1] > 200 and got[3] > 25000


def test_product_space_closed_forms(gpu):
    import torch
    from edlgt_amd.device import SectorBasis
    print("SRE2 closed forms on 6 qubits", flush=True)
    B = SectorBasis(np.full(6, 2, dtype=np.int32))
    assert B.n_states == 64
    cfg = B.configs().astype(np.int64)
    assert np.array_equal(cfg, np.array(list(itertools.product(range(2), repeat=6))))
    ts = [np.pi / 8, 0.3, np.pi / 4, 0.1, 1.1, np.pi / 8]
    psi = np.array([1.0 + 0j])
    want = 0.0
    for t in ts:
        psi = np.kron(psi, np.array([np.cos(t), np.sin(t)]))
        a, b = np.cos(t) ** 2, np.sin(t) ** 2
        want -= np.log((a * a + b * b) ** 2 + 12 * a * a * b * b)
    got = B.sre2(torch.as_tensor(psi, device="cuda"), 1e-12)
    assert got[1] == 64 and got[3] == 64 and abs(got[0] - want) < 1e-12
    e = np.zeros(64, complex)
    e[37] = 1.0
    assert abs(B.sre2(e, 1e-2)[0]) < 1e-15
    u = np.full(64, 1 / 8, complex)
    assert abs(B.sre2(u, 1e-9)[0]) < 1e-13
