"""Pins the CPU oracle against the reference's own known-answer tests
(validation/qed/test_qed_1d_matter.py:8-72, test_qed_2d_matter.py:8-74,
test_qed_2d_pure.py:8-71: per-site ground-state energies at atol 1e-10) and
the independent transfer-matrix sector counts."""
import numpy as np
import pytest

from oracle.model import QEDOracle, lowest_eigs

KATS = [
    # lvals, has_obc, pure, g, m, E0 per site, sector dim
    ([4], [True], False, 1.0, 1.0, -0.5691411466075982, 6),
    ([4], [False], False, 1.0, 1.0, -0.5903881075118848, 13),
    ([4, 2], [True, True], False, 1.0, 1.0, -0.6862839217038432, None),
    ([3, 3], [True, True], True, 1.0, None, -0.09443358823578024, None),
    ([3, 3], [False, False], True, 1.0, None, -0.2042469295595554, None),
]


@pytest.mark.parametrize("lvals,obc,pure,g,m,e0,dim", KATS)
def test_qed_kat(lvals, obc, pure, g, m, e0, dim):
    M = QEDOracle(1, pure, lvals, obc)
    if dim is not None:
        assert M.N == dim
    H = M.hamiltonian(g, m)
    e = lowest_eigs(H, 1)[0] / np.prod(lvals)
    assert abs(e - e0) < 1e-10


def test_qed_2d_pbc_kat_and_c2_dimension():
    M = QEDOracle(1, False, [4, 2], [False, False])
    assert M.N == 140647
    H = M.hamiltonian(1.0, 1.0)
    e = lowest_eigs(H, 1)[0] / 8
    assert abs(e - (-0.8830087027925408)) < 1e-10


def test_c1_dimension():
    assert QEDOracle(1, False, [10], [True]).N == 197


def test_qed_3d_theta_kat():
    """validation/qed/test_qed_3d_theta.py: 2x2x2 pure PBC (flux through the
    three coordinate planes at the origin fixed to 0), g=2.3, theta=0.41."""
    M = QEDOracle(1, True, [2, 2, 2], [False, False, False])
    e = lowest_eigs(M.hamiltonian(2.3, theta=0.41), 1)[0] / 8
    assert abs(e - (-0.37555632787741716)) < 1e-10
