"""The C/OpenMP restatement (oracle/refcpu.c, the CPU baseline) against the
Python oracle: CSR SpMV against SciPy, sector enumeration against the oracle's
sector tables, two-pass n-body triplets against the oracle's triplets (same
rows, columns, values and emission order)."""
import numpy as np
import pytest

from oracle import lattice as lat
from oracle import refcpu
from oracle.model import QEDOracle, SU2Oracle, ZnOracle, nbody_triplets


@pytest.fixture(scope="module")
def qed():
    return QEDOracle(1, False, [4, 2], [True, True])


def test_spmv_matches_scipy(qed):
    H = qed.hamiltonian(1.0, 0.7)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(H.shape[0]) + 1j * rng.standard_normal(H.shape[0])
    y = np.empty_like(x)
    refcpu.spmv(H.indptr.astype(np.int64), H.indices.astype(np.int32), H.data, x, y)
    ref = H @ x
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    Z = ZnOracle(3, [2, 2], [False, False], (0, 0)).hamiltonian(1.0, 0.5)
    xr = rng.standard_normal(Z.shape[0])
    yr = np.empty_like(xr)
    refcpu.spmv(Z.indptr.astype(np.int64), Z.indices.astype(np.int32), Z.data, xr, yr)
    assert np.max(np.abs(yr - Z @ xr)) <= 1e-13 * np.max(np.abs(Z @ xr))


def _dg(orc, name):
    return np.real(np.einsum("sii->si", orc.ops[name]))


def test_sector_matches_oracle(qed):
    pairs = lat.link_pairs(qed.lvals, qed.has_obc)
    link = np.stack([np.stack([_dg(qed, f"E_p{d}"), -_dg(qed, f"E_m{d}")]) for d in qed.dirs])
    cfg = refcpu.sector(qed.loc_dims, _dg(qed, "N")[None], np.array([qed.n_sites // 2], float), link,
                        np.zeros(2), pairs)
    assert np.array_equal(cfg, qed.configs)
    su = SU2Oracle(False, [8], [True], sectors=[8])
    pairs = lat.link_pairs(su.lvals, su.has_obc)
    link = np.stack([np.stack([_dg(su, "T2_px"), -_dg(su, "T2_mx")])])
    cfg = refcpu.sector(su.loc_dims, _dg(su, "N_tot")[None], np.array([8.0]), link, np.zeros(1), pairs)
    assert np.array_equal(cfg, su.configs)
    zn = ZnOracle(3, [3, 2], [False, False], None)
    pairs = lat.link_pairs(zn.lvals, zn.has_obc)
    link = np.stack([np.stack([_dg(zn, f"E_p{d}"), _dg(zn, f"E_m{d}")]) for d in "xy"])
    cfg = refcpu.sector(zn.loc_dims, link_diags=link, link_sectors=np.zeros(2), pairs=pairs)
    assert np.array_equal(cfg, zn.configs)


def test_nbody_triplets_match_oracle(qed):
    names = ["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"]
    stack = np.stack([qed.ops[n] for n in names])
    plaq = lat.plaquette_sites(lat.zig_zag(qed.lvals, 0), qed.lvals, ["x", "y"], qed.has_obc)
    hop = np.stack([qed.ops["Q_px_dag"], qed.ops["Q_mx"]])
    for st, sites in ((stack, plaq), (hop, [0, 1]), (hop, [5, 6])):
        r, c, v, miss = refcpu.nbody(qed.configs, st, sites)
        ro, co, vo = nbody_triplets(st, sites, qed.configs, qed.keys, qed.loc_dims)
        assert miss == 0
        assert np.array_equal(r, ro) and np.array_equal(c, co)
        assert np.max(np.abs(v - vo), initial=0.0) == 0.0


def test_cpu_build_and_e0_matches_kat_and_oracle():
    """The host baseline's full path (sector, triplets, CSR, eigsh) on the
    reference's 4x2 PBC KAT (validation/qed/test_qed_2d_matter.py:43-74)."""
    e0, N, nnz, T = refcpu.qed_build_and_e0([4, 2], [False, False], 1.0, 1.0)
    assert N == 140647
    assert nnz == QEDOracle(1, False, [4, 2], [False, False]).hamiltonian(1.0, 1.0).nnz
    assert abs(e0 / 8 - (-0.8830087027925408)) < 1e-10
