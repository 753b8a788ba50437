"""Momentum-sector projector of the reference (oracle restatement; test
infrastructure).

Follows edlgt/symmetries/translational_sym.py: block translations as site
permutations in the module's own row-major coordinates (last axis fastest,
:131-157, :508-565), the translation table by lookup of every translated
configuration (build_all_translations, :568-618), one reference per orbit
(its smallest basis index) with its axis periods and the momentum filter
(k_d p_d) % R_d == 0 (select_references, :621-695), and the Bloch column
(1/sqrt(norm)) sum over the period box of exp(-2 pi i k.t / R) T(t)|ref>, equal
images accumulated, amplitudes |a| <= 1e-14 dropped for k != 0, rows sorted
(momentum_basis_zero_k / _finite_k, :698-950, helpers :220-413).
The translation-combined (TC_symmetry) variant of the 1D SU(2) chain is
restated in momentum_basis_tc (build_TC_translations :111-128,
momentum_basis_zero_k_TC / _finite_k_TC :457-505, 953-1306).
"""
import numpy as np
import scipy.sparse as sp


def source_sites(lvals, unit_cell):
    """source[s, dest] = site whose state lands on dest under block shift s
    (_prepare_translation_source_sites)."""
    lvals = np.asarray(lvals, dtype=np.int64)
    uc = np.asarray(unit_cell, dtype=np.int64)
    R = lvals // uc
    D = len(lvals)
    n = int(np.prod(lvals))
    out = np.empty((int(np.prod(R)), n), dtype=np.int64)
    for s in range(out.shape[0]):
        t, f = np.empty(D, dtype=np.int64), s
        for ax in range(D - 1, -1, -1):
            t[ax] = f % R[ax]
            f //= R[ax]
        for src in range(n):
            c, q = np.empty(D, dtype=np.int64), src
            for ax in range(D - 1, -1, -1):
                c[ax] = q % lvals[ax]
                q //= lvals[ax]
            c = (c + t * uc) % lvals
            dest = 0
            for ax in range(D):
                dest = dest * lvals[ax] + c[ax]
            out[s, dest] = src
    return out, R


def translation_table(configs, keys, loc_dims, lvals, unit_cell):
    """trans[i, s] = basis index of block shift s applied to configuration i."""
    from .model import encode
    src, R = source_sites(lvals, unit_cell)
    trans = np.empty((configs.shape[0], src.shape[0]), dtype=np.int64)
    for s in range(src.shape[0]):
        rolled = configs[:, src[s]]
        k = encode(rolled, loc_dims)
        idx = np.searchsorted(keys, k)
        if np.any(idx >= len(keys)) or np.any(keys[np.minimum(idx, len(keys) - 1)] != k):
            raise ValueError("the sector is not closed under the translations")
        trans[:, s] = idx
    return trans, R


def momentum_basis(configs, keys, loc_dims, lvals, unit_cell, k_vals):
    """B (n_rows x n_cols, CSC, complex128): column c = normalised Bloch sum of
    reference c; columns in increasing reference index."""
    trans, R = translation_table(configs, keys, loc_dims, lvals, unit_cell)
    D = len(R)
    strides = np.ones(D, dtype=np.int64)
    for ax in range(D - 2, -1, -1):
        strides[ax] = strides[ax + 1] * R[ax + 1]
    k_vals = np.asarray(k_vals, dtype=np.int64)
    finite = bool(np.any(k_vals != 0))
    # per-axis phase tables by repeated multiplication (_prepare_phase_table)
    ptab = np.ones((D, int(R.max())), dtype=np.complex128)
    for ax in range(D):
        if R[ax] > 1:
            step = np.exp(-1j * 2.0 * np.pi * (k_vals[ax] % R[ax]) / float(R[ax]))
            for q in range(1, R[ax]):
                ptab[ax, q] = ptab[ax, q - 1] * step
    N = configs.shape[0]
    refs = np.nonzero(trans.min(axis=1) == np.arange(N))[0]
    cols_r, cols_c, cols_v = [], [], []
    n_cols = 0
    for ref in refs:
        row = trans[ref]
        p = np.empty(D, dtype=np.int64)
        for ax in range(D):
            p[ax] = R[ax]
            for q in range(1, R[ax] + 1):
                if row[(q % R[ax]) * strides[ax]] == ref:
                    p[ax] = q
                    break
        if np.any((k_vals * p) % R != 0):
            continue
        acc = {}
        order = []
        for flat in range(int(np.prod(p))):
            t, f = np.empty(D, dtype=np.int64), flat
            for ax in range(D - 1, -1, -1):
                t[ax] = f % p[ax]
                f //= p[ax]
            img = int(row[int(np.dot(t, strides))])
            ph = 1.0
            if finite:
                ph = 1.0 + 0.0j
                for ax in range(D):
                    ph *= ptab[ax, t[ax]]
            if img not in acc:
                acc[img] = 0.0
                order.append(img)
            acc[img] += ph
        rows = np.array(order, dtype=np.int64)
        vals = np.array([acc[i] for i in order], dtype=np.complex128)
        if finite:
            keep = np.abs(vals) > 1e-14
            rows, vals = rows[keep], vals[keep]
            nrm = np.sum(np.abs(vals) ** 2)
            if nrm <= 1e-30 or rows.size == 0:
                n_cols += 1  # the column exists, empty (reference keeps it)
                continue
        else:
            nrm = np.sum(np.abs(vals) ** 2)
        o = np.argsort(rows)
        cols_r.append(rows[o])
        cols_c.append(np.full(rows.size, n_cols))
        cols_v.append(vals[o] / np.sqrt(nrm))
        n_cols += 1
    if cols_r:
        r, c, v = np.concatenate(cols_r), np.concatenate(cols_c), np.concatenate(cols_v)
    else:
        r = c = np.zeros(0, dtype=np.int64)
        v = np.zeros(0, dtype=np.complex128)
    return sp.csc_matrix((v, (r, c)), shape=(N, n_cols))


# Translation-combined (TC) symmetry of the 1D SU(2) chain with staggered
# matter: X = T o C (the local map C, then a shift by one site), restated from
# translational_sym.py:457-505 (build_TC_translations), :111-128
# (precompute_C_sign_per_config), :953-1122 (momentum_basis_zero_k_TC) and
# :1123-1306 (momentum_basis_finite_k_TC).
SU2_TC_MAP = np.array([4, 5, 2, 3, 0, 1])
SU2_TC_SIGN = np.array([1.0, 1.0, 1.0, -1.0, 1.0, 1.0])


def tc_translation_table(configs, keys, loc_dims, c_map=SU2_TC_MAP):
    """trans[i, t] = index of X^t |config_i> (X^t: C applied when t is odd,
    then a shift by t sites)."""
    from .model import encode
    N, L = configs.shape
    trans = np.empty((N, L), dtype=np.int64)
    for t in range(L):
        cur = configs if t % 2 == 0 else c_map[configs]
        rolled = np.empty_like(cur)
        rolled[:, (np.arange(L) + t) % L] = cur
        k = encode(rolled, loc_dims)
        idx = np.searchsorted(keys, k)
        if np.any(idx >= N) or np.any(keys[np.minimum(idx, N - 1)] != k):
            raise ValueError("the sector is not closed under the TC translations")
        trans[:, t] = idx
    return trans


def momentum_basis_tc(configs, keys, loc_dims, k, c_map=SU2_TC_MAP, c_sign=SU2_TC_SIGN):
    """B for TC momentum k (1D): reference = smallest index of its X-orbit,
    period p = smallest p > 0 with X^p ref = ref, kept iff (k p) % L == 0;
    column sum over t < p of exp(-2 pi i k t / L) s_t X^t|ref>, s_t = 1 for even
    t and the reference's C sign for odd t; k = 0 drops columns whose norm^2 is
    within 1e-8 of zero (np.isclose), k != 0 drops |a| <= 1e-14 and keeps
    cancelled columns empty."""
    trans = tc_translation_table(configs, keys, loc_dims, c_map)
    N, L = configs.shape
    sign = np.prod(c_sign[configs], axis=1)
    refs = np.nonzero(trans.min(axis=1) == np.arange(N))[0]
    finite = (k % L) != 0 or k != 0
    rr, cc, vv = [], [], []
    n_cols = 0
    for ref in refs:
        row = trans[ref]
        p = L
        for q in range(1, L + 1):
            if row[q % L] == ref:
                p = q
                break
        if (k * p) % L != 0:
            continue
        acc, order = {}, []
        for t in range(p):
            incr = sign[ref] if t % 2 == 1 else 1.0
            ph = np.exp(-1j * 2.0 * np.pi * ((k % L) * t) / float(L)) if finite else 1.0
            img = int(row[t])
            if img not in acc:
                acc[img] = 0.0
                order.append(img)
            acc[img] += incr * ph
        rows = np.array(order, dtype=np.int64)
        vals = np.array([acc[i] for i in order], dtype=np.complex128)
        if finite:
            keep = np.abs(vals) > 1e-14
            rows, vals = rows[keep], vals[keep]
            nrm = float(np.sum(np.abs(vals) ** 2))
            if nrm <= 1e-30 or rows.size == 0:
                n_cols += 1
                continue
        else:
            nrm = float(np.sum(np.abs(vals) ** 2))
            if np.isclose(nrm, 0.0):
                n_cols += 1
                continue
        o = np.argsort(rows)
        rr.append(rows[o])
        cc.append(np.full(rows.size, n_cols))
        vv.append(vals[o] / np.sqrt(nrm))
        n_cols += 1
    if rr:
        r, c, v = np.concatenate(rr), np.concatenate(cc), np.concatenate(vv)
    else:
        r = c = np.zeros(0, dtype=np.int64)
        v = np.zeros(0, dtype=np.complex128)
    return sp.csc_matrix((v, (r, c)), shape=(N, n_cols))
