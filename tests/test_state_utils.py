"""CPU: QMB_state's post-processing helpers (qmb_state.py:79-157, 357-391):
normalize, truncate, expectation_value and the participation Renyi entropy,
checked on host tensors against closed forms: Pauli expectation values of a
known qubit state in the three operator formats, Renyi entropies of a basis
state (0), of a uniform superposition over N states (log N for every alpha),
and of a two-state superposition (log(p^a + q^a) / (1 - a)); plus the
reference's argument errors."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from edlgt_amd.models import QMB_state


def _state(vec):
    return QMB_state(torch.as_tensor(np.asarray(vec, dtype=np.complex128)), [len(vec)], [len(vec)])


def test_normalize_and_truncate():
    s = _state([3.0, 4.0j, 1e-15])
    assert abs(s.normalize() - 5.0) < 1e-14
    assert abs(s.get_norm() - 1.0) < 1e-15
    assert np.allclose(s.psi, [0.6, 0.8j, 2e-16])
    assert abs(s.normalize() - 1.0) < 1e-15  # already normalised: unchanged
    t = s.truncate(1e-14)
    assert t[2] == 0 and t[0] == s.psi[0] and t[1] == s.psi[1] and s.psi[2] != 0
    with pytest.raises(ValueError):
        s.truncate(-1.0)


def test_pauli_expectation_values_in_every_format():
    th, ph = 0.37, 1.1
    # |psi> = cos(th)|0> + e^{i ph} sin(th)|1>: <X> = sin 2th cos ph, <Y> = sin 2th sin ph, <Z> = cos 2th
    s = _state([np.cos(th), np.exp(1j * ph) * np.sin(th)])
    X = np.array([[0, 1], [1, 0]], complex)
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1.0, -1.0]).astype(complex)
    want = {"X": np.sin(2 * th) * np.cos(ph), "Y": np.sin(2 * th) * np.sin(ph), "Z": np.cos(2 * th)}
    for name, M in (("X", X), ("Y", Y), ("Z", Z)):
        r, c = np.nonzero(M)
        for op in (M, sp.csr_matrix(M), sp.coo_matrix(M), (r, c, M[r, c])):
            assert abs(s.expectation_value(op) - want[name]) < 1e-15
            assert abs(s.expectation_value(op, "imag")) < 1e-15
    # a non-Hermitian operator has an imaginary part: <psi|i Z|psi> = i cos 2th
    assert abs(s.expectation_value(1j * Z, "imag") - np.cos(2 * th)) < 1e-15
    with pytest.raises(TypeError):
        s.expectation_value([[1, 0], [0, 1]])
    with pytest.raises(ValueError):
        s.expectation_value(np.eye(3))
    with pytest.raises(ValueError):
        s.expectation_value(Z, "abs")


@pytest.mark.parametrize("alpha", [0.5, 2, 3])
def test_participation_renyi_entropy_closed_forms(alpha):
    assert abs(_state([0, 1, 0, 0]).participation_renyi_entropy(alpha)) < 1e-15
    n = 64
    u = _state(np.full(n, 1 / np.sqrt(n)) * np.exp(1j * np.arange(n)))
    assert abs(u.participation_renyi_entropy(alpha) - np.log(n)) < 1e-12
    p = 0.3
    w = _state([np.sqrt(p), 1j * np.sqrt(1 - p)])
    assert abs(w.participation_renyi_entropy(alpha) - np.log(p ** alpha + (1 - p) ** alpha) / (1 - alpha)) < 1e-14


def test_participation_renyi_entropy_rejects_alpha():
    for a in (0, 1, -2):
        with pytest.raises(ValueError):
            _state([1, 0]).participation_renyi_entropy(a)


def test_entropy_partition_half():
    """get_entropy_partition (qmb_hamiltonian.py:635-664): the first half of a
    chain, or the first half of each of the first two rows of a 2D lattice."""
    from edlgt_amd.models import get_entropy_partition
    assert get_entropy_partition([6]) == [0, 1, 2]
    assert get_entropy_partition([5]) == [0, 1]
    assert get_entropy_partition([4, 3]) == [0, 1, 4, 5]
    with pytest.raises(NotImplementedError):
        get_entropy_partition([4], option="quarter")


def test_density_matrix_projector_and_mixed_expectation():
    """get_projector_for_efficient_density_matrix (qmb_state.py:549-592): the
    leading eigenvectors above the threshold, relaxed by tens until two pass;
    mixed_exp_val_data (:657-679): <psi1|O|psi2> from triplets equals the
    dense product, repeated triplets summed."""
    from edlgt_amd.models import get_projector_for_efficient_density_matrix, mixed_exp_val_data
    rng = np.random.default_rng(5)
    vals = np.array([0.05, 0.7, 0.2, 0.05])
    Q, _ = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))
    P = get_projector_for_efficient_density_matrix(vals, Q, 0.1)
    assert P.shape == (4, 2) and np.allclose(P, Q[:, [1, 2]])
    P = get_projector_for_efficient_density_matrix(vals, Q, 0.5)  # one passes: relaxed to 0.05
    assert P.shape == (4, 2)
    P = get_projector_for_efficient_density_matrix(np.array([0.9, 0.01, 0.005]), np.eye(3), 0.5)
    assert P.shape == (3, 2)  # 0.5 -> 0.05 -> 0.005
    r = np.array([0, 1, 2, 2, 3])
    c = np.array([1, 1, 0, 0, 3])
    v = np.array([1.0, 2j, 0.5, 0.25, -1.0])
    O = np.zeros((4, 4), complex)
    np.add.at(O, (r, c), v)
    a = rng.normal(size=4) + 1j * rng.normal(size=4)
    b = rng.normal(size=4) + 1j * rng.normal(size=4)
    assert abs(mixed_exp_val_data(a, b, r, c, v) - np.vdot(a, O @ b)) < 1e-14


def test_block_ids_projection_and_loc_dims_printout(caplog):
    """build_energy_block_ids (quantum_model.py:1451-1478), apply_projection
    (:1411-1448) and format_loc_dims (:1481-1523) on closed cases."""
    import logging
    import edlgt_amd.models as M
    b, n = M.build_energy_block_ids(np.array([-2.0, -2.0 + 1e-12, -1.0, 0.5, 0.5]), 1e-10)
    assert list(b) == [0, 0, 1, 2, 2] and n == 3
    b, n = M.build_energy_block_ids(np.array([1.0]))
    assert list(b) == [0] and n == 1
    P = np.array([[1.0, 0.0], [0.0, 1j], [0.0, 0.0]])
    O = np.arange(9.0).reshape(3, 3)
    want = P.conj().T @ O @ P
    assert np.allclose(M.apply_projection(P, O), want)
    got = M.apply_projection(sp.csr_matrix(P), sp.csr_matrix(O))
    assert sp.issparse(got) and np.allclose(got.toarray(), want)
    with pytest.raises(TypeError):
        M.apply_projection(sp.csr_matrix(P), O)
    with pytest.raises(ValueError):
        M.apply_projection(P[:2], O)
    with caplog.at_level(logging.INFO, logger=M.logger.name):
        M.format_loc_dims(np.arange(12), [2, 3, 2], pad=2)
    lines = [r.getMessage() for r in caplog.records]
    assert lines == ["[ 0  1]   [ 6  7]", "[ 2  3]   [ 8  9]", "[ 4  5]   [10 11]"]


def test_truncation_norm_and_sorted_indices():
    """truncation, get_norm, get_sorted_indices (qmb_state.py:466-523)."""
    import edlgt_amd.models as M
    a = np.array([1e-15, -0.5, 2e-14j, 0.3 + 0.4j])
    assert np.array_equal(M.truncation(a), np.array([0, -0.5, 2e-14j, 0.3 + 0.4j]))
    assert np.array_equal(M.truncation(a, 1e-13), np.array([0, -0.5, 0, 0.3 + 0.4j]))
    assert M.get_norm(np.array([3.0, 4j])) == 5.0
    # |.| = 0.5 three times: ties by real part, largest first; then 0.2
    d = np.array([0.2, -0.5, 0.5, 0.5j])
    assert list(M.get_sorted_indices(d)) == [2, 3, 1, 0]


def test_exact_evolution_overlaps_and_dtype_modes():
    """exact_time_evolution / compute_overlap_with_eigenstates
    (qmb_hamiltonian.py:792-861) against scipy's expm on a random Hermitian
    matrix; the dtype-mode labels (quantum_model.py:1342-1409)."""
    from scipy.linalg import expm
    import edlgt_amd.models as M
    rng = np.random.default_rng(2)
    A = rng.normal(size=(7, 7)) + 1j * rng.normal(size=(7, 7))
    H = A + A.conj().T
    e, V = np.linalg.eigh(H)
    psi = rng.normal(size=7) + 1j * rng.normal(size=7)
    tl = np.array([0.0, 0.3, 1.7])
    out = M.exact_time_evolution(tl, e, V, psi)
    for k, t in enumerate(tl):
        assert np.max(np.abs(out[k] - expm(-1j * t * H) @ psi)) < 1e-12
    assert np.max(np.abs(M.compute_overlap_with_eigenstates(V, psi) - V.conj().T @ psi)) < 1e-14
    norm = M.QuantumModel._normalize_dtype_mode
    assert norm(True) == "complex" and norm(False) == "real" and norm(None) == "auto" and norm(" Real ") == "real"
    with pytest.raises(ValueError):
        norm(None, allow_auto=False)
    with pytest.raises(ValueError):
        norm("half")
    with pytest.raises(TypeError):
        norm(3)
    qm = object.__new__(M.QuantumModel)
    assert qm.configure_dtype_mode() == "complex" and qm.configure_dtype_mode("auto", "real") == "real"
    assert qm.configure_dtype_mode(False) == "real" and qm.dtype_mode == "real"
