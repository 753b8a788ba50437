"""Dictionary-coded CSR values (edlgt_csr_compress).

A coded product multiplies by exactly the stored values in the same order as
the uncoded kernels, so SpMV and the fused coefficient SpMM give the same bits
coded and uncoded (where both take the same kernel shape); the grouped
products of large coded matrices are checked against the uncoded product at
1e-12 relative and for batch independence (bitwise).  Eigenvalues of coded
matrices are checked against the oracle."""
import numpy as np
import pytest

from oracle.model import QEDOracle, ZnOracle, lowest_eigs

pytestmark = pytest.mark.gpu


def _qed(lvals, obc):
    from edlgt_amd import QED_Model
    return QED_Model(spin=1, pure_theory=False, lvals=lvals, has_obc=obc)


def test_coded_spmv_bitwise_complex(gpu):
    import torch
    mod = _qed([4, 2], [False, False])
    H = mod.build_Hamiltonian(1.0, 0.5)
    rng = np.random.default_rng(11)
    x = torch.as_tensor(rng.standard_normal(H.n) + 1j * rng.standard_normal(H.n), device=gpu)
    X = torch.as_tensor(rng.standard_normal((3, H.n)) + 1j * rng.standard_normal((3, H.n)), device=gpu)
    y0 = H.matvec(x).cpu().numpy()
    Y0 = H.matvec(X).cpu().numpy()
    nd = H.compress()
    print(f"QED 4x2 n_dict={nd} code_bytes={H.code_bytes}", flush=True)
    assert 0 < nd <= 4096
    y1 = H.matvec(x).cpu().numpy()
    assert np.array_equal(y0, y1)
    Y1 = H.matvec(X).cpu().numpy()
    # block products: coded == uncoded bitwise, a column's bits independent of
    # the block; single-vector products (another kernel) agree to rounding
    assert np.array_equal(Y0, Y1)
    assert np.array_equal(H.matvec(X[1:]).cpu().numpy(), Y1[1:])
    for j in range(3):
        yj = H.matvec(X[j].contiguous()).cpu().numpy()
        assert np.max(np.abs(Y1[j] - yj)) <= 1e-13 * np.max(np.abs(yj))
    ref = H.to_scipy() @ X.cpu().numpy().T
    assert np.max(np.abs(Y1 - ref.T)) <= 1e-13 * np.max(np.abs(ref))
    # eigenvalues of the coded matrix against the oracle
    orc = QEDOracle(1, False, [4, 2], [False, False])
    ev, _, _, _ = H.lanczos(2, tol=1e-12, want_vectors=False)
    assert np.max(np.abs(ev - lowest_eigs(orc.hamiltonian(1.0, 0.5), 2))) < 1e-10


def test_coded_coef_spmm_small_bitwise_and_server(gpu):
    import torch
    from edlgt_amd import Server, Zn_Model
    orc = ZnOracle(3, [3, 3], [False, False], (0, 0))
    mod = Zn_Model(3, lvals=[3, 3], has_obc=[False, False], flux_sector=(0, 0))
    Hu = mod.build_multi_term()
    Hc = mod.build_multi_term()
    nd = Hc.compress()
    assert 0 < nd <= 256 and Hc.code_bytes == 1
    rng = np.random.default_rng(3)
    X = torch.as_tensor(rng.standard_normal((7, Hu.n)), device=gpu)
    coef = np.stack([rng.uniform(-2, 2, 7), rng.uniform(0.1, 3, 7)], axis=1)
    assert np.array_equal(Hu.matvec_coef(X, coef).cpu().numpy(), Hc.matvec_coef(X, coef).cpu().numpy())
    reqs = [(i, [1.0, h], 40 + i) for i, h in enumerate([0.5, 1.5, 2.5])]
    a = Server(Hu, n_eigs=3, tol=1e-10, max_basis=24, max_slots=3).solve_all(reqs)
    b = Server(Hc, n_eigs=3, tol=1e-10, max_basis=24, max_slots=3).solve_all(reqs)
    for rid, coef_r, _ in reqs:
        assert np.array_equal(a[rid][0], b[rid][0])
        ref = lowest_eigs(orc.hamiltonian(*coef_r), 3)
        assert np.max(np.abs(b[rid][0] - ref)) < 1e-9 * np.max(np.abs(ref))


def test_coded_coef_spmm_large_per_column(gpu):
    """4x4 zero-flux sector (14.3M rows): coded coefficient products run in
    zero-padded groups of 4 packed columns; a column's bits do not depend on
    the batch it is in."""
    import torch
    from edlgt_amd import Zn_Model
    mod = Zn_Model(3, lvals=[4, 4], has_obc=[False, False], flux_sector=(0, 0))
    Hu = mod.build_multi_term()
    rng = np.random.default_rng(5)
    X = torch.as_tensor(rng.standard_normal((3, Hu.n)), device=gpu)
    coef = np.array([[1.0, 0.4], [1.0, 2.9], [-0.5, 1.3]])
    Yu = Hu.matvec_coef(X, coef).cpu().numpy()
    nd = Hu.compress()
    print(f"Z3 4x4 n_dict={nd}", flush=True)
    assert 0 < nd <= 4096
    Yc = Hu.matvec_coef(X, coef).cpu().numpy()
    assert np.max(np.abs(Yc - Yu)) <= 1e-12 * np.max(np.abs(Yu))
    for j in (0, 2):
        y1 = Hu.matvec_coef(X[j:j + 1].contiguous(), coef[j:j + 1]).cpu().numpy()
        assert np.array_equal(y1[0], Yc[j])
    y5 = Hu.matvec_coef(torch.cat([X, X[:2]]).contiguous(), np.vstack([coef, coef[:2]])).cpu().numpy()
    assert np.array_equal(y5[:3], Yc) and np.array_equal(y5[3:], Yc[:2])
    # all-ones coefficients agree with the plain SpMV (another kernel, another
    # summation order per row)
    y_all = Hu.matvec(X[0].contiguous()).cpu().numpy()
    y_one = Hu.matvec_coef(X[0:1].contiguous(), [[1.0, 1.0]]).cpu().numpy()[0]
    assert np.max(np.abs(y_all - y_one)) <= 1e-13 * np.max(np.abs(y_all))
