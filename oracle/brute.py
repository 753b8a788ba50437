"""Full-space Kronecker Hamiltonians for tiny lattices (oracle; test infrastructure).

Restates the reference's brute-force cross-check
(validation/qed/test_qed_1d_integrated.py:13-152): build H in the full product
of per-site gauge-invariant bases with sparse Kronecker products
(qmb_operations.py:28-248), then restrict it to the sector rows/columns with
the expansion projector (generate_configs.py:365-389).  Independent of the
sector triplet builder: it never looks up a configuration.
"""
from math import prod

import numpy as np
import scipy.sparse as sp

from . import lattice as lat


def _embed(ops_by_site, loc_dims):
    """kron over sites of the given per-site matrices (identity where missing)."""
    out = None
    for s, d in enumerate(loc_dims):
        m = ops_by_site.get(s)
        m = sp.identity(d, format="csr") if m is None else sp.csr_matrix(m[:d, :d])
        out = m if out is None else sp.kron(out, m, format="csr")
    return out


def qed_full_hamiltonian(model, g, m=None):
    """Same term list as QEDOracle.triplets, in the full product space."""
    c = model.couplings(g, m)
    ld = model.loc_dims
    H = 0
    for s in range(model.n_sites):
        H = H + c["E"] * _embed({s: model.ops["E2"][s]}, ld)
    if model.dim > 1:
        names = ["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"]
        for i in range(model.n_sites):
            p = lat.plaquette_sites(lat.zig_zag(model.lvals, i), model.lvals, ["x", "y"], model.has_obc)
            if p is None:
                continue
            if len(set(p)) < 4:
                raise NotImplementedError("plaquette touching a site twice")
            T = c["B"] * _embed({s: model.ops[n][s] for s, n in zip(p, names)}, ld)
            H = H + T + T.conj().T
    if not model.pure:
        for i in range(model.n_sites):
            co = lat.zig_zag(model.lvals, i)
            stag = "even" if sum(co) % 2 == 0 else "odd"
            H = H + c[f"m_{stag}"] * _embed({i: model.ops["N"][i]}, ld)
        for d in model.dirs:
            for i in range(model.n_sites):
                co = lat.zig_zag(model.lvals, i)
                b = lat.neighbor_sites(co, model.lvals, d, model.has_obc)
                if b is None:
                    continue
                if b[0] == b[1]:
                    raise NotImplementedError("bond from a site to itself")
                stag = "even" if sum(co) % 2 == 0 else "odd"
                T = c[f"t{d}_{stag}"] * _embed({b[0]: model.ops[f"Q_p{d}_dag"][b[0]],
                                                b[1]: model.ops[f"Q_m{d}"][b[1]]}, ld)
                H = H + T + T.conj().T
    return sp.csr_matrix(H)


def restrict_to_sector(H_full, configs, loc_dims):
    """P^T H P with P the 0/1 expansion projector of the sector rows."""
    st = np.ones(len(loc_dims), dtype=np.int64)
    for i in range(len(loc_dims) - 2, -1, -1):
        st[i] = st[i + 1] * int(loc_dims[i + 1])
    idx = (configs.astype(np.int64) * st).sum(axis=1)
    return H_full[idx][:, idx]


def zn_full_hamiltonian(model, J, h):
    """Full-space Z_n Hamiltonian (same terms as ZnOracle.hamiltonian)."""
    ld = model.loc_dims
    H = 0
    for s in range(model.n_sites):
        H = H + (-h) * _embed({s: model.ops["E_clock"][s]}, ld)
    names = ["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"]
    for i in range(model.n_sites):
        p = lat.plaquette_sites(lat.zig_zag(model.lvals, i), model.lvals, ["x", "y"], model.has_obc)
        if p is None:
            continue
        T = (-J) * _embed({s: model.ops[n][s] for s, n in zip(p, names)}, ld)
        H = H + T + T.conj().T
    return sp.csr_matrix(H)
