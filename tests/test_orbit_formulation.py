"""CPU: the per-state orbit formulation that the device orbit scan
(`orbit_scan` / `orbit_fill` in csrc/momentum.hip) computes, restated in numpy
and checked against the oracle's momentum basis, which follows the
reference's period-box construction (translational_sym.py:621-950).

Each state j reads only its own translation-table row: representative m =
row minimum, stabiliser H = {s : T_s j = j}, rank in the orbit
#{s : T_s j < j} / |H|, amplitude sum over {s : T_s j = m} of
exp(+2 pi i k.s / R) divided by sqrt(S |H|); a kept orbit on which k is not
trivial on H gives an empty column.  That this equals the box sum (same
columns, same row order, amplitudes to rounding) is a statement about the
translation group, not a restatement of the oracle's loop, so a wrong sign of
the phase, a wrong rank or a dropped empty column fails here."""
import itertools

import numpy as np
import pytest

from oracle.model import SU2Oracle, ZnOracle
from oracle.momentum import momentum_basis, translation_table


def _orbit_formulation(configs, keys, loc_dims, lvals, unit_cell, k_vals):
    trans, R = translation_table(configs, keys, loc_dims, lvals, unit_cell)
    N, S = trans.shape
    R = [int(r) for r in R]
    D = len(R)
    sstride = [int(np.prod(R[d + 1:])) for d in range(D)]
    kd = [int(k_vals[d]) % R[d] for d in range(D)]
    s_all = np.arange(S)
    sd = np.stack([(s_all // sstride[d]) % R[d] for d in range(D)])  # (D, S)
    # exp(-2 pi i k.s / R) == 1, exactly on the integers
    k_triv = (sum(kd[d] * sd[d] * (S // R[d]) for d in range(D)) % S) == 0
    # per-axis phase tables by repeated multiplication; phase of shift -s
    ptab = np.ones((D, max(R)), dtype=np.complex128)
    for d in range(D):
        step = np.exp(-2j * np.pi * kd[d] / R[d])
        for q in range(1, R[d]):
            ptab[d, q] = ptab[d, q - 1] * step
    cph = np.ones(S, dtype=np.complex128)
    for d in range(D):
        cph = cph * ptab[d, (R[d] - sd[d]) % R[d]]
    idx = np.arange(N)
    m = trans.min(axis=1)
    rep = m == idx
    keep = rep.copy()
    for d in range(D):
        p = np.full(N, R[d])
        for q in range(R[d], 0, -1):  # smallest q wins
            p = np.where(trans[:, (q % R[d]) * sstride[d]] == idx, q, p)
        keep &= (kd[d] * p) % R[d] == 0
    fix = trans == idx[:, None]
    h = fix.sum(axis=1)
    triv = np.all(~fix | k_triv[None, :], axis=1)
    cnt = np.where(keep, np.where(triv, S // h, 0), 0)
    col_of = np.concatenate([[0], np.cumsum(keep)])
    first = np.concatenate([[0], np.cumsum(cnt)])
    n_cols, nnz = int(col_of[-1]), int(first[-1])
    col_ptr = np.full(n_cols + 1, -1, dtype=np.int64)
    col_ptr[n_cols] = nnz
    col_ptr[col_of[idx[keep]]] = first[idx[keep]]
    rows = np.full(nnz, -1, dtype=np.int64)
    vals = np.zeros(nnz, dtype=np.complex128)
    live = keep[m] & (cnt[m] > 0)
    amp = ((trans == m[:, None]) * cph[None, :]).sum(axis=1)
    less = (trans < idx[:, None]).sum(axis=1)
    pos = first[m] + less // h
    assert len(np.unique(pos[live])) == nnz
    rows[pos[live]] = idx[live]
    vals[pos[live]] = amp[live] / np.sqrt(S * h[live])
    assert np.all(rows >= 0) and np.all(col_ptr >= 0)
    return col_ptr, rows, vals, n_cols


def _check(configs, keys, loc_dims, lvals, unit_cell, k):
    col_ptr, rows, vals, n_cols = _orbit_formulation(configs, keys, loc_dims, lvals, unit_cell, k)
    Bo = momentum_basis(configs, keys, loc_dims, lvals, unit_cell, k)
    Bo.sort_indices()
    assert Bo.shape[1] == n_cols
    assert np.array_equal(Bo.indptr, col_ptr) and np.array_equal(Bo.indices, rows)
    assert np.max(np.abs(Bo.data - vals), initial=0.0) < 1e-15


@pytest.mark.parametrize("k", [0, 1, 2])
def test_su2_chain_unit_cell_two(k):
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    _check(orc.configs, orc.keys, orc.loc_dims, [6], [2], [k])


@pytest.mark.parametrize("k", [0, 1, 3])
def test_su2_chain_unit_cell_one(k):
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    _check(orc.configs, orc.keys, orc.loc_dims, [6], [1], [k])


def test_z3_square_all_momenta():
    z = ZnOracle(3, [3, 3], [False, False], flux_sector=(0, 0))
    for kx, ky in itertools.product(range(3), range(3)):
        _check(z.configs, z.keys, z.loc_dims, [3, 3], [1, 1], [kx, ky])
