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


@pytest.mark.parametrize("dim,pure", [(1, False), (2, False), (2, True), (3, False)])
def test_qed_background_tables_equal_oracle(dim, pure):
    """Static charges q in {-1, 0, 1}: the engine's per-charge labels hold
    exactly the oracle's states of that charge (the reference keeps, per site,
    the columns of its charge out of a basis with the charge as leading dof)
    and the same operators, atol 0."""
    site = opx.qed_site(1, pure, dim, backgrounds={-1, 0, 1})
    basis, states = qed_ops.gauge_invariant_states(1, pure, dim, background=1)
    ops_o = qed_ops.dressed_site_operators(1, pure, dim, background=1)
    names = ["E2", "N", "bg", "E_px"] + (["C_px,py", "C_my,px"] if dim >= 2 else []) + \
        ([] if pure else [f"Q_p{a}_dag" for a in "xyz"[:dim]])
    names = [n for n in names if n in ops_o]
    for lab, P in basis.items():
        for q in (-1, 0, 1):
            cols = np.flatnonzero(states[lab][:, 0] == q)
            key = opx.bg_label(lab, q)
            if cols.size == 0:
                assert key not in site.states
                continue
            assert np.array_equal(site.values[key], states[lab][cols, 1:])
            mine = opx.qed_operators(site, 1, pure, dim, key, with_bg=True)
            Pq = P[:, cols]
            for name in names:
                assert np.array_equal(mine[name], (Pq.T @ ops_o[name] @ Pq).toarray()), (key, name)


def test_qed_named_states_equal_oracle_physics():
    """QED_Model.overlap_QMB_state's 1D spin-1 index tables (QED_model.py:796-846)
    name the states the oracle finds from their physical content.  The
    meson's odd site must be a bulk site: on a 4-site open chain the table
    puts it on the edge, where index 1 is the filled vacuum and the pair is
    not a sector state (the reference's table does the same)."""
    from types import SimpleNamespace

    from edlgt_amd.models import QED_Model
    from oracle.model import QEDOracle
    for n in (4, 10):
        for obc in (True, False):
            orc = QEDOracle(1, False, [n], [obc])
            stub = SimpleNamespace(dim=1, pure_theory=False, spin=1, n_sites=n, has_obc=[obc], bg_list=None)
            for name in ("V", "PV", "meson_center"):
                if name == "meson_center" and obc and n < 6:
                    continue
                mine = QED_Model.overlap_QMB_state(stub, name)
                ref = orc.label_config("meson_center" if name == "meson_center" else "V")
                assert np.array_equal(mine, ref), (n, obc, name)
    stub = SimpleNamespace(dim=2, pure_theory=False, spin=1, n_sites=4, has_obc=[True, True], bg_list=None)
    with pytest.raises(NotImplementedError):
        QED_Model.overlap_QMB_state(stub, "V")
