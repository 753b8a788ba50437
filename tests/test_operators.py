"""Engine-side dressed-site tables vs the oracle's Kronecker-then-project
construction (QED_operators.py:120-538, quantum_model.py:400-502): equal with
atol 0 on every label, plus the algebraic invariants the reference asserts."""
import numpy as np
import pytest

from edlgt_amd import operators as opx
from oracle import qed_ops


@pytest.mark.parametrize("dim,pure", [(1, False), (2, False), (2, True), (1, True)])
def test_qed_tables_equal_oracle(dim, pure):
    site = opx.qed_site(1, pure, dim)
    basis, states = qed_ops.gauge_invariant_states(1, pure, dim)
    ops_o = qed_ops.dressed_site_operators(1, pure, dim)
    for lab, P in basis.items():
        assert site.dim(lab) == P.shape[1]
        mine = opx.qed_operators(site, 1, pure, dim, lab)
        for name, O in ops_o.items():
            ref = (P.T @ O @ P).toarray()
            assert np.array_equal(mine[name], ref), (lab, name)


def test_bulk_dimensions():
    # d_loc of the reference's bulk labels: 1D matter 5, 2D matter 35, 2D pure 19
    assert opx.qed_site(1, False, 1).dim("even") == 5
    assert opx.qed_site(1, False, 2).dim("even") == 35
    assert opx.qed_site(1, True, 2).dim("site") == 19


def test_gauss_law_and_isometry():
    basis, _ = qed_ops.gauge_invariant_states(1, False, 2)
    ops = qed_ops.dressed_site_operators(1, False, 2)
    from scipy.sparse import identity
    for core, off in (("even", 0), ("odd", 1)):
        P = basis[core]
        assert abs((P.T @ P - identity(P.shape[1])).sum()) == 0
        G = ops["E_px"] + ops["E_py"] - ops["E_mx"] - ops["E_my"] - ops["N"] + off * identity(ops["N"].shape[0])
        assert abs(G @ P).max() == 0


def test_corner_and_hopping_fermionic_signs():
    # hopping across a bond is gauge invariant: Q_px_dag at i times Q_mx at j
    # preserves sum_x (E_px - E_mx) per site pair; corners commute (Zn_operators.py:262-265)
    ops = qed_ops.dressed_site_operators(1, False, 2)
    C1, C2 = ops["C_px,py"], ops["C_mx,my"]
    assert abs(C1 @ C2 - C2 @ C1).max() < 1e-15
