"""Stabilizer Renyi-2 entropy on a truncated support (oracle restatement;
test infrastructure).

Follows edlgt/modeling/qmb_state.py: QMB_state.stabilizer_renyi_entropy
(:393-465) and extract_support (:708-790), with the kernels of
edlgt/tools/magic.py: mixed-radix keys with the rightmost site fastest
(compute_strides :41-76), the X-strings p = (beta - alpha) mod d of every
ordered pair of support configurations (all_pairwise_pkeys_support
:283-352), and per string the exact contribution T_p = sum_kappa |V_kappa|^2,
V_kappa = sum over ordered pairs (a, b) with alpha_a + alpha_b = kappa (mod d)
of A_a conj(A_b), A_alpha = c_alpha conj(c_{alpha+p}) over the alpha whose
shift stays in the support (exact_Xstring_from_support :555-649).  A loop over strings, as the reference
runs them, each string's pairs grouped by kappa in numpy.
"""
import numpy as np


def strides(loc_dims):
    d = np.asarray(loc_dims, dtype=np.int64)
    s = np.ones(d.size, dtype=np.int64)
    for k in range(d.size - 2, -1, -1):
        s[k] = s[k + 1] * d[k + 1]
    return s


def extract_support(psi, loc_dims, configs, delta=1e-2):
    """(indices, coeffs, configs, keys, discarded) sorted by key: the smallest
    set of largest-probability configurations holding at least 1 - delta."""
    if not 0.0 < delta < 1.0:
        raise ValueError("prob_threshold must be 1> delta >0.")
    prob = np.abs(psi) ** 2
    order = np.argsort(-prob, kind="mergesort")
    cum = np.cumsum(prob[order], dtype=np.float64)
    k_last = int(np.searchsorted(cum, 1.0 - delta, side="left"))
    idx = order[: k_last + 1].astype(np.int64)
    coeffs = psi[idx]
    cfg = np.asarray(configs)[idx].astype(np.int64)
    discarded = max(0.0, 1.0 - float(np.sum(np.abs(coeffs) ** 2)))
    keys = cfg @ strides(loc_dims)
    o = np.argsort(keys, kind="mergesort")
    return idx[o], coeffs[o], cfg[o], keys[o], discarded


def sre2(psi, loc_dims, configs, delta=1e-2):
    """(SRE2, number of X-strings, support size, discarded weight)."""
    d = np.asarray(loc_dims, dtype=np.int64)
    st = strides(d)
    _, c, cfg, keys, discarded = extract_support(np.asarray(psi, dtype=np.complex128), d, configs, delta)
    c = c / np.sqrt(float(np.sum(np.abs(c) ** 2)))
    K = c.size
    pk = (((cfg[None, :, :] - cfg[:, None, :]) % d) @ st).ravel()
    strings = np.unique(pk)
    M2 = 0.0
    for p in strings:
        pv = (int(p) // st) % d
        sk = ((cfg + pv) % d) @ st
        j = np.searchsorted(keys, sk)
        j = np.minimum(j, K - 1)
        hit = keys[j] == sk
        ia = np.nonzero(hit)[0]
        A = c[ia] * np.conj(c[j[hit]])
        kap = (((cfg[ia][:, None, :] + cfg[ia][None, :, :]) % d) @ st).ravel()
        val = (A[:, None] * np.conj(A[None, :])).ravel()
        _, inv = np.unique(kap, return_inverse=True)
        V = np.bincount(inv, weights=val.real) + 1j * np.bincount(inv, weights=val.imag)
        M2 += float(np.sum(np.abs(V) ** 2))
    return -float(np.log(max(M2, 1e-16))), strings.size, K, discarded
