"""Host layout of the device Schmidt kernels (models.schmidt_layout): the
packed per-component blocks hold exactly the reference's psi_matrix entries
(qmb_state.py:160-209), so the blocks' spectra are the dense SVD spectrum
and their M M^H blocks, placed by row_ids, are the dense reduced density
matrix (qmb_state.py:211-243).  Checked against the oracle's dense
restatement on SU(2) and QED sectors, no GPU."""
import numpy as np
import pytest

from oracle.entropy import dense_rdm, dense_schmidt_weights, entropy, partition_maps
from oracle.model import QEDOracle, SU2Oracle


def _blocks(psi, lay):
    M = np.zeros(lay["m_elems"], dtype=np.complex128)
    M[lay["pos_host"]] = psi
    for d, c, moff, goff in lay["blocks"]:
        yield M[moff:moff + d * c].reshape(d, c), goff


CASES = [
    ("su2", [0, 1, 2, 3, 4]), ("su2", [0, 2, 5]), ("su2", [9]), ("su2", [3, 4, 5, 6]),
    ("qed", [0, 1, 2, 3]), ("qed", [0, 4]), ("qed", [5]),
]


@pytest.fixture(scope="module")
def systems():
    su2 = SU2Oracle(False, [10], [True], sectors=[10])
    qed = QEDOracle(1, False, [4, 2], [True, True])
    return {"su2": su2.configs, "qed": qed.configs}


@pytest.mark.parametrize("name,keep", CASES)
def test_layout_reproduces_dense_psi_matrix(systems, name, keep):
    from edlgt_amd.models import schmidt_layout
    configs = systems[name]
    rng = np.random.default_rng(len(keep) + 7)
    psi = rng.standard_normal(len(configs)) + 1j * rng.standard_normal(len(configs))
    psi /= np.linalg.norm(psi)
    part = partition_maps(configs, keep)
    ref = np.sort(dense_schmidt_weights(psi, part))[::-1]
    small = schmidt_layout(part, "small")
    pos = small["pos_host"]
    assert np.unique(pos).size == pos.size and pos.min() >= 0 and pos.max() < small["m_elems"]
    d, c = small["blocks"][:, 0], small["blocks"][:, 1]
    assert np.all(d <= c) and small["n_vals"] == d.sum()
    lam = np.concatenate([np.linalg.eigvalsh(M @ M.conj().T) for M, _ in _blocks(psi, small)])
    lam = np.sort(lam)[::-1]
    n = min(lam.size, ref.size)
    assert np.max(np.abs(lam[:n] - ref[:n])) < 1e-13
    assert np.all(np.abs(lam[n:]) < 1e-13) and np.all(ref[n:] < 1e-13)
    assert abs(entropy(lam) - entropy(ref)) < 1e-12
    sub = schmidt_layout(part, "subsys")
    rho = np.zeros((part["subsys_dim"],) * 2, dtype=np.complex128)
    for (M, _), rows in zip(_blocks(psi, sub), sub["row_ids"]):
        rho[np.ix_(rows, rows)] = M @ M.conj().T
    assert np.max(np.abs(rho - dense_rdm(psi, part))) < 1e-14
    # every subsystem configuration is a row of exactly one block
    assert np.array_equal(np.sort(np.concatenate(sub["row_ids"])), np.arange(part["subsys_dim"]))
    assert schmidt_layout(part, "small") is small  # cached


def test_layout_rejects_unknown_orientation():
    from edlgt_amd.models import schmidt_layout
    part = partition_maps(np.array([[0, 1], [1, 0]]), [0])
    with pytest.raises(ValueError):
        schmidt_layout(part, "columns")
