"""GPU tests of the batched path: multi-term assembly (H(J, h) = J H_J + h H_h
on one pattern) against the oracle's single-coupling matrix, the fused
per-column-coefficient SpMM against the oracle product and against itself one
column at a time (bitwise), and the continuous-batching server: eigenvalues
against the oracle, and every request's result bitwise the same alone and in
a mix with staggered admission."""
import numpy as np
import pytest

from oracle.model import ZnOracle, lowest_eigs

pytestmark = pytest.mark.gpu


def _zn(lv, fs=(0, 0)):
    from edlgt_amd import Zn_Model
    return ZnOracle(3, lv, [False, False], fs), Zn_Model(3, lvals=lv, has_obc=[False, False], flux_sector=fs)


@pytest.mark.parametrize("lv", [[2, 2], [3, 3]])
def test_multi_term_equals_single_build(gpu, lv):
    print(f"multi-term lvals={lv}", flush=True)
    orc, mod = _zn(lv)
    H2 = mod.build_multi_term()
    assert H2.n_terms == 2
    t = H2.terms()
    assert set(np.unique(t)) <= {0, 1}
    rp, col, val = H2.to_host()
    # inside a row: columns ascending, one entry per (col, term), terms ascending
    for r in range(0, H2.n, max(1, H2.n // 200)):
        c, tt = col[rp[r]:rp[r + 1]], t[rp[r]:rp[r + 1]]
        key = c.astype(np.int64) * 2 + tt
        assert np.all(np.diff(key) > 0)
    for J, h in [(1.0, 0.8), (0.3, 2.5), (-1.2, 0.0)]:
        Ho = orc.hamiltonian(J, h)
        Hm = H2.to_scipy(coef=[J, h])
        d = abs(Hm - Ho)
        assert (d.max() if d.nnz else 0.0) <= 1e-14 * max(1.0, abs(Ho).max())


def test_coef_spmm_parity_and_column_independence(gpu):
    import torch
    orc, mod = _zn([3, 3])
    H2 = mod.build_multi_term()
    rng = np.random.default_rng(7)
    nb = 11
    X = rng.standard_normal((nb, H2.n))
    coef = np.stack([rng.uniform(-2, 2, nb), rng.uniform(0.1, 3, nb)], axis=1)
    Xd = torch.tensor(X, device=gpu)
    Y = H2.matvec_coef(Xd, coef).cpu().numpy()
    for j in range(nb):
        ref = orc.hamiltonian(*coef[j]) @ X[j]
        assert np.max(np.abs(Y[j] - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    # each column alone gives the same bits as inside the batch of 11
    for j in (0, 5, 10):
        y1 = H2.matvec_coef(Xd[j:j + 1].contiguous(), coef[j:j + 1]).cpu().numpy()
        assert np.array_equal(y1[0], Y[j])
    # complex vectors on the real matrix
    Z = torch.tensor(X[:3] + 1j * X[3:6], device=gpu)
    Yz = H2.matvec_coef(Z, coef[:3]).cpu().numpy()
    for j in range(3):
        ref = orc.hamiltonian(*coef[j]) @ (X[j] + 1j * X[3 + j])
        assert np.max(np.abs(Yz[j] - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_server_eigenvalues_and_alone_vs_mix(gpu):
    from edlgt_amd import Server
    orc, mod = _zn([3, 3])
    H2 = mod.build_multi_term()
    k = 3
    reqs = [(i, [1.0, h], 1000 + i) for i, h in enumerate([0.4, 0.9, 1.7, 2.6, 0.6, 1.2])]
    srv = Server(H2, n_eigs=k, tol=1e-10, max_basis=24, max_slots=4)
    mix = srv.solve_all(reqs)
    assert sorted(mix) == [r[0] for r in reqs]
    for rid, coef, _ in reqs:
        ev, es, mv, conv = mix[rid]
        assert conv and mv > 0
        ref = lowest_eigs(orc.hamiltonian(*coef), k)
        assert np.max(np.abs(ev - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))
    # the same requests alone, each in a fresh single-slot server
    for rid, coef, seed in (reqs[0], reqs[3], reqs[5]):
        alone = Server(H2, n_eigs=k, tol=1e-10, max_basis=24, max_slots=1).solve_all([(rid, coef, seed)])
        assert np.array_equal(alone[rid][0], mix[rid][0])
        assert np.array_equal(alone[rid][1], mix[rid][1])
        assert alone[rid][2] == mix[rid][2]
    # and in another mix (different neighbours and admission order)
    mix2 = Server(H2, n_eigs=k, tol=1e-10, max_basis=24, max_slots=3).solve_all(reqs[::-1])
    for rid in mix:
        assert np.array_equal(mix2[rid][0], mix[rid][0])


def test_server_complex_vectors_and_full_slots(gpu):
    from edlgt_amd import Server
    orc, mod = _zn([2, 2], fs=None)
    H2 = mod.build_multi_term()
    srv = Server(H2, n_eigs=2, tol=1e-10, max_basis=16, max_slots=2, complex_vectors=True)
    assert srv.submit(0, [1.0, 0.5], 1) and srv.submit(1, [0.5, 1.0], 2)
    assert not srv.submit(2, [1.0, 1.0], 3)  # both slots busy
    assert srv.active == 2
    out = {}
    while srv.active:
        for rid, ev, es, mv, cv in srv.step():
            out[rid] = ev
    for rid, coef in [(0, [1.0, 0.5]), (1, [0.5, 1.0])]:
        ref = lowest_eigs(orc.hamiltonian(*coef), 2)
        assert np.max(np.abs(out[rid] - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))


def test_server_warm_start_matches_cold(gpu):
    from edlgt_amd import Server
    orc, mod = _zn([3, 3])
    H2 = mod.build_multi_term()
    reqs = [(i, [1.0, float(h)], 500 + i) for i, h in enumerate(np.linspace(0.4, 2.8, 9))]
    cold = Server(H2, n_eigs=3, tol=1e-10, max_basis=24, max_slots=3).solve_all(reqs)
    srv = Server(H2, n_eigs=3, tol=1e-10, max_basis=24, max_slots=3, warm_cache=4)
    warm = srv.solve_all(reqs)
    for rid, coef, _ in reqs:
        ref = lowest_eigs(orc.hamiltonian(*coef), 3)
        assert warm[rid][3]
        assert np.max(np.abs(warm[rid][0] - cold[rid][0])) < 1e-8 * max(1.0, np.max(np.abs(ref)))
        assert np.max(np.abs(warm[rid][0] - ref)) < 1e-8 * max(1.0, np.max(np.abs(ref)))
    assert sum(v[2] for v in warm.values()) < sum(v[2] for v in cold.values())
    inf = srv.info()
    assert inf["active"] == 0 and inf["allocated"] == 3
    assert inf["shared_bytes"] >= 4 * 3 * H2.n * 8
    srv.trim()
    assert srv.info()["allocated"] == 0
    # a trimmed server allocates again on demand
    again = srv.solve_all([(99, [1.0, 1.1], 7)])
    ref = lowest_eigs(orc.hamiltonian(1.0, 1.1), 3)
    assert np.max(np.abs(again[99][0] - ref)) < 1e-8 * max(1.0, np.max(np.abs(ref)))


def test_qed_c2_g_sweep_multi_term(gpu):
    """QED 4x2 PBC with matter (N = 140,647), m = 0.5, g in {0.5, 1, 2}: one
    multi-term assembly; at every g the summed matrix has the oracle's pattern
    and values, and the served lowest 4 eigenvalues match the oracle."""
    from edlgt_amd import QED_Model, Server
    from oracle.model import QEDOracle
    orc = QEDOracle(1, False, [4, 2], [False, False])
    mod = QED_Model(spin=1, pure_theory=False, lvals=[4, 2], has_obc=[False, False])
    H3 = mod.build_multi_term(0.5)
    assert H3.n_terms == 3 and H3.n == 140647
    gs = [0.5, 1.0, 2.0]
    for g in gs:
        Ho = orc.hamiltonian(g, 0.5)
        Hm = H3.to_scipy(coef=mod.multi_term_coefficients(g))
        assert np.array_equal(Hm.indptr, Ho.indptr) and np.array_equal(Hm.indices, Ho.indices)
        assert np.max(np.abs(Hm.data - Ho.data)) <= 1e-14 * np.max(np.abs(Ho.data))
    srv = Server(H3, n_eigs=4, tol=1e-10, max_basis=64, max_slots=3)
    res = srv.solve_all([(i, mod.multi_term_coefficients(g), 7 + i) for i, g in enumerate(gs)])
    for i, g in enumerate(gs):
        ref = lowest_eigs(orc.hamiltonian(g, 0.5), 4)
        assert np.max(np.abs(res[i][0] - ref)) < 1e-9


def test_slot_bytes_match_device_allocation(gpu):
    """edlgt_server_info's slot_bytes is the device memory one slot takes: the
    first submission allocates it (free memory drops by slot_bytes, within the
    allocator's rounding); the Gram result buffer is pinned host memory."""
    import torch
    from edlgt_amd import Server
    _, mod = _zn([4, 3])
    H2 = mod.build_multi_term()
    srv = Server(H2, n_eigs=4, tol=1e-8, max_basis=64, max_slots=2)
    info = srv.info()
    assert info["allocated"] == 0
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    assert srv.submit(0, [1.0, 0.9], 5)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert srv.info()["allocated"] == 1
    used = free0 - free1
    assert abs(used - info["slot_bytes"]) <= 4 << 20, (used, info["slot_bytes"])
    srv.solve_all([])


@pytest.mark.parametrize("coded", [False, True])
def test_server_graph_replay_equals_eager(gpu, coded, monkeypatch):
    """The fused product of a served step is captured once per column count
    into a HIP graph and replayed with the step's pointers and coefficients
    from pinned staging memory; eager launches (EDLGT_SERVER_GRAPHS=0) give
    the same results bit for bit.  Six slots of 3 columns: 18 columns in two
    chunks, then fewer as requests finish, each count replayed many times
    with the solvers' own kernels and synchronisations between replays.
    (k = 3: the next level of Z3 3x3 is four-fold degenerate.)"""
    import torch
    from edlgt_amd import Server
    orc, mod = _zn([3, 3])
    H2 = mod.build_multi_term()
    if coded:
        assert H2.compress() > 0
    reqs = [(i, [1.0, h], 300 + i) for i, h in enumerate([0.5, 0.8, 1.3, 1.9, 2.4, 0.7, 2.9])]
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("EDLGT_SERVER_GRAPHS", mode)
        srv = Server(H2, n_eigs=3, tol=1e-10, max_basis=32, max_slots=6, stream=torch.cuda.Stream())
        out[mode] = srv.solve_all(reqs)
        del srv
    for rid, coef, _ in reqs:
        assert np.array_equal(out["1"][rid][0], out["0"][rid][0])
        assert np.array_equal(out["1"][rid][1], out["0"][rid][1])
        assert out["1"][rid][2] == out["0"][rid][2]
        ref = lowest_eigs(orc.hamiltonian(*coef), 3)
        assert np.max(np.abs(out["1"][rid][0] - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))
