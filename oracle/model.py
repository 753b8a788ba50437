"""Sector enumeration, triplet assembly, CSR and solvers (oracle; test infrastructure).

Restatement of:
  * QuantumModel.project_operators          models/quantum_model.py:400-502
  * get_symmetry_sector_generators          symmetries/symmetry_sector.py:148-196
  * iterative_sitebased_sym_sector_configs  symmetries/symmetry_sector.py:785-967
  * iterative_link_sector_configs_plus      symmetries/symmetry_sector.py:1329-1469
  * _prepare_local_transition_data          symmetries/sym_qmb_operations.py:32-92
  * nbody_data / _2sites / _4sites          symmetries/sym_qmb_operations.py:297-562
  * localbody_data_par                      symmetries/sym_qmb_operations.py:565-602
  * Local/TwoBody/PlaquetteTerm drivers     modeling/{local,twobody,plaquette}_term.py
  * QED_Model.__init__/build_truncated_Hamiltonian/QED_Hamiltonian_couplings
                                            models/QED_model.py:23-394, 719-776
  * QMB_hamiltonian.build("sparse")/diagonalize
                                            modeling/qmb_hamiltonian.py:173-332

The configuration lookup uses mixed-radix int64 keys (site 0 most
significant) and ``np.searchsorted``; key order equals the lexicographic row
order the reference's binary search relies on (generate_configs.py:145-173,
243-264; the reference itself uses this equivalence in
quantum_model.py:73-117).  A lookup miss raises (the reference writes -1).
"""
from math import prod

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import eigsh, expm_multiply

from . import lattice as lat
from . import qed_ops

OP_TOL = 1e-10   # sym_qmb_operations.py:21
SYM_TOL = 1e-13  # symmetry_sector.py:17


# ----------------------------------------------------------------------------
# keys
# ----------------------------------------------------------------------------
def strides(loc_dims):
    """Mixed-radix strides, last site fastest (tools/magic.py:40-80)."""
    st = np.ones(len(loc_dims), dtype=np.int64)
    for i in range(len(loc_dims) - 2, -1, -1):
        st[i] = st[i + 1] * int(loc_dims[i + 1])
    return st


def encode(configs, loc_dims):
    return (np.asarray(configs, dtype=np.int64) * strides(loc_dims)[None, :]).sum(axis=1)


# ----------------------------------------------------------------------------
# projection
# ----------------------------------------------------------------------------
def project_operators(ops, basis, labels, dtype=np.complex128, site_cols=None):
    """P^T O P per site, zero-padded to d_max (quantum_model.py:400-502);
    site_cols[j] (optional) keeps only those columns of site j's isometry (the
    states of its background charge, quantum_model.py:440-466)."""
    proj = [basis[l] if site_cols is None else basis[l][:, site_cols[j]] for j, l in enumerate(labels)]
    loc_dims = np.array([P.shape[1] for P in proj], dtype=np.int64)
    dmax = int(loc_dims.max())
    out = {}
    for name, O in ops.items():
        arr = np.zeros((len(labels), dmax, dmax), dtype=dtype)
        for j, l in enumerate(labels):
            P = proj[j]
            eff = (P.conj().T @ O @ P).toarray()
            arr[j, : loc_dims[j], : loc_dims[j]] = eff
        out[name] = arr
    return out, loc_dims


# ----------------------------------------------------------------------------
# sector enumeration
# ----------------------------------------------------------------------------
def sector_configs(loc_dims, glob_diags=None, glob_sectors=None, glob_type="U",
                   link_diags=None, link_sectors=None, pairs=None,
                   nbody_diags=None, nbody_sectors=None, nbody_sites=None,
                   nbody_type="U"):
    """Level-by-level prefix expansion with activated-constraint pruning.

    Candidates at level k are prev*d_k + s (prefix-major), filtered in order,
    so the result is lexicographically sorted (symmetry_sector.py:785-967,
    1329-1469).  glob_diags: (n_glob, n_sites, dmax); link_diags:
    (n_dirs, 2, n_sites, dmax) with pairs[d] an (n_pairs, 2) array;
    nbody_diags: (n_nb, n_sites, dmax) with nbody_sites[k] the site list."""
    n = len(loc_dims)
    n_glob = 0 if glob_diags is None else glob_diags.shape[0]
    if n_glob:
        rem_min = np.zeros((n_glob, n + 1))
        rem_max = np.zeros((n_glob, n + 1))
        for g in range(n_glob):
            for s in range(n - 1, -1, -1):
                v = glob_diags[g, s, : loc_dims[s]]
                rem_min[g, s] = rem_min[g, s + 1] + v.min()
                rem_max[g, s] = rem_max[g, s + 1] + v.max()
    # pairs activated at the max endpoint
    act_pairs = [[] for _ in range(n)]
    if link_diags is not None:
        for d, pl in enumerate(pairs):
            for (i, j) in pl:
                act_pairs[max(i, j)].append((d, int(i), int(j)))
    act_nb = [[] for _ in range(n)]
    if nbody_diags is not None:
        for k, sites in enumerate(nbody_sites):
            act_nb[max(sites)].append(k)
    configs = np.zeros((1, 0), dtype=np.int64)
    gvals = np.zeros((1, n_glob))
    for s in range(n):
        d = int(loc_dims[s])
        nprev = configs.shape[0]
        cand = np.empty((nprev * d, s + 1), dtype=np.int64)
        cand[:, :s] = np.repeat(configs, d, axis=0)
        cand[:, s] = np.tile(np.arange(d), nprev)
        ok = np.ones(nprev * d, dtype=bool)
        gnew = np.repeat(gvals, d, axis=0)
        for g in range(n_glob):
            if glob_type == "U":
                gnew[:, g] += glob_diags[g, s, cand[:, s]]
                if s < n - 1:
                    lo = gnew[:, g] + rem_min[g, s + 1]
                    hi = gnew[:, g] + rem_max[g, s + 1]
                    ok &= ~((glob_sectors[g] < lo - SYM_TOL) | (glob_sectors[g] > hi + SYM_TOL))
                else:
                    ok &= np.abs(gnew[:, g] - glob_sectors[g]) <= SYM_TOL
            else:
                gnew[:, g] = (gnew[:, g] if s > 0 else 1.0) * glob_diags[g, s, cand[:, s]]
                if s == n - 1:
                    ok &= np.abs(gnew[:, g] - glob_sectors[g]) <= SYM_TOL
        for (dd, i, j) in act_pairs[s]:
            v = link_diags[dd, 0, i, cand[:, i]] + link_diags[dd, 1, j, cand[:, j]]
            ok &= np.abs(v - link_sectors[dd]) <= SYM_TOL
        for k in act_nb[s]:
            if nbody_type == "Z":
                v = np.ones(cand.shape[0])
                for site in nbody_sites[k]:
                    v *= nbody_diags[k, site, cand[:, site]]
            else:
                v = np.zeros(cand.shape[0])
                for site in nbody_sites[k]:
                    v += nbody_diags[k, site, cand[:, site]]
            if nbody_type.startswith("M"):  # additive modulo m (Z_m flux)
                m = int(nbody_type[1:])
                ok &= (np.rint(v).astype(np.int64) - int(round(nbody_sectors[k]))) % m == 0
            else:
                ok &= np.abs(v - nbody_sectors[k]) <= SYM_TOL
        configs = cand[ok]
        gvals = gnew[ok]
    return configs.astype(np.uint16)


def sector_configs_exhaustive(loc_dims, check):
    """Exhaustive filter over the full product space (sitebased_sym_sector_configs,
    symmetry_sector.py:27-65); ``check(config)->bool``.  Tiny lattices only."""
    from itertools import product
    out = [c for c in product(*[range(int(d)) for d in loc_dims]) if check(c)]
    return np.array(out, dtype=np.uint16).reshape(-1, len(loc_dims))


# ----------------------------------------------------------------------------
# triplets
# ----------------------------------------------------------------------------
def transitions(op_stack, sites):
    """_prepare_local_transition_data (sym_qmb_operations.py:32-92)."""
    d = op_stack.shape[-1]
    tabs = []
    for f, s in enumerate(sites):
        O = op_stack[f, s]
        kets = [[b for b in range(d) if abs(O[a, b]) > OP_TOL] for a in range(d)]
        vals = [[O[a, b] for b in kets[a]] for a in range(d)]
        tabs.append((kets, vals))
    return tabs


def nbody_triplets(op_stack, sites, configs, keys, loc_dims):
    """(row, col, val) of an n-body operator in the sector basis.

    n_ops == 1 -> diagonal only, keep |O_aa| >= OP_TOL (localbody_data_par,
    sym_qmb_operations.py:565-602); n_ops >= 2 -> Cartesian product of local
    transitions, amplitude = ordered product of factors, column found by key
    lookup (nbody_data*, :297-562)."""
    N = configs.shape[0]
    sites = [int(s) for s in sites]
    if len(sites) == 1:
        s = sites[0]
        a = configs[:, s].astype(np.int64)
        v = op_stack[0, s][a, a]
        keep = np.abs(v) >= OP_TOL
        r = np.nonzero(keep)[0]
        return r, r, v[keep].astype(np.complex128)
    st = strides(loc_dims)
    tabs = transitions(op_stack, sites)
    nf = len(sites)
    maxc = [max(len(k) for k in t[0]) for t in tabs]
    rows_out, cols_out, vals_out = [], [], []
    from itertools import product
    a_loc = [configs[:, s].astype(np.int64) for s in sites]
    for combo in product(*[range(m) for m in maxc]):
        valid = np.ones(N, dtype=bool)
        dkey = np.zeros(N, dtype=np.int64)
        amp = np.ones(N, dtype=np.complex128)
        for f in range(nf):
            kets, vals = tabs[f]
            d = len(kets)
            cnt = np.array([len(kets[a]) for a in range(d)])
            ket_tab = np.array([kets[a][combo[f]] if combo[f] < len(kets[a]) else -1 for a in range(d)])
            val_tab = np.array([vals[a][combo[f]] if combo[f] < len(kets[a]) else 0 for a in range(d)],
                               dtype=np.complex128)
            valid &= cnt[a_loc[f]] > combo[f]
            kt = ket_tab[a_loc[f]]
            dkey += np.where(valid, (kt - a_loc[f]) * st[sites[f]], 0)
            amp = amp * val_tab[a_loc[f]]
        r = np.nonzero(valid)[0]
        if r.size == 0:
            continue
        tk = keys[r] + dkey[r]
        c = np.searchsorted(keys, tk)
        c = np.minimum(c, N - 1)
        if not np.all(keys[c] == tk):
            raise ValueError("ket configuration not found in the symmetry sector")
        rows_out.append(r)
        cols_out.append(c)
        vals_out.append(amp[r])
    if not rows_out:
        z = np.zeros(0, dtype=np.int64)
        return z, z, np.zeros(0, dtype=np.complex128)
    return np.concatenate(rows_out), np.concatenate(cols_out), np.concatenate(vals_out)


def expval_nbody(op_list, sites, configs, keys, loc_dims, psi):
    """<psi|O|psi> of one operator product on concrete sites: exp_val_data
    (qmb_state.py:630-653) over the sector triplets of nbody_term
    (sym_qmb_operations.py:167-602)."""
    r, c, v = nbody_triplets(np.stack(op_list), sites, configs, keys, loc_dims)
    return complex(np.sum(np.conj(psi[r]) * v * psi[c]))


def twobody_corr(op1, op2, configs, keys, loc_dims, psi):
    """lattice_twobody_exp_val (twobody_term.py:232-264): corr[i, j] =
    corr[j, i] = Re <psi|O1_i O2_j|psi> for i < j, zero diagonal."""
    n = configs.shape[1]
    corr = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            corr[i, j] = corr[j, i] = expval_nbody([op1, op2], [i, j], configs, keys, loc_dims, psi).real
    return corr


class Triplets:
    """Accumulator mirroring QMB_hamiltonian.add_term/build (qmb_hamiltonian.py:155-225)."""

    def __init__(self, N):
        self.N = N
        self.r, self.c, self.v = [], [], []

    def add(self, r, c, v):
        self.r.append(np.asarray(r, dtype=np.int64))
        self.c.append(np.asarray(c, dtype=np.int64))
        self.v.append(np.asarray(v, dtype=np.complex128))

    def csr(self, real=False):
        """Canonical CSR: duplicates summed, columns sorted, explicit zeros kept
        (SciPy coalescing used by ham_format='sparse', qmb_hamiltonian.py:206-209)."""
        r = np.concatenate(self.r)
        c = np.concatenate(self.c)
        v = np.concatenate(self.v)
        if real:
            if np.max(np.abs(v.imag), initial=0.0) > 1e-12:
                raise ValueError("complex values in real mode")
            v = v.real.copy()
        H = sp.coo_matrix((v, (r, c)), shape=(self.N, self.N)).tocsr()
        H.sum_duplicates()
        H.sort_indices()
        return H


def add_local(T, op_stack, configs, keys, loc_dims, strength, mask_sites=None):
    """LocalTerm.get_Hamiltonian in the sector (local_term.py:55-117)."""
    rs, cs, vs = [], [], []
    for s in range(configs.shape[1]):
        if mask_sites is not None and not mask_sites[s]:
            continue
        r, c, v = nbody_triplets(op_stack[None], [s], configs, keys, loc_dims)
        rs.append(r); cs.append(c); vs.append(v)
    T.add(np.concatenate(rs), np.concatenate(cs), np.concatenate(vs) * strength)


def add_nbody_list(T, op_list, site_lists, configs, keys, loc_dims, strength, add_dagger):
    """TwoBodyTerm / PlaquetteTerm sector branch (twobody_term.py:116-158,
    plaquette_term.py:116-163): concatenate all placements, scale, then append
    the dagger block (col, row, conj(val))."""
    stack = np.stack(op_list)
    rs, cs, vs = [], [], []
    for sites in site_lists:
        r, c, v = nbody_triplets(stack, sites, configs, keys, loc_dims)
        rs.append(r); cs.append(c); vs.append(v)
    r = np.concatenate(rs); c = np.concatenate(cs); v = np.concatenate(vs) * strength
    T.add(r, c, v)
    if add_dagger:
        T.add(c, r, np.conj(v))


# ----------------------------------------------------------------------------
# QED model
# ----------------------------------------------------------------------------
class QEDOracle:
    """QED_Model (truncated spin, ladder U, no background) restated
    (models/QED_model.py:23-233, 235-394, 719-776)."""

    def __init__(self, spin, pure_theory, lvals, has_obc, enumerate_sector=True, bg_list=None,
                 link_symmetries=True):
        self.spin, self.pure = spin, pure_theory
        # link_symmetries=False: no link constraints in the sector, penalty
        # terms in H instead (QED_model.py:131-141, 381-393)
        self.link_symmetries = bool(link_symmetries)
        self.lvals, self.has_obc = list(lvals), list(has_obc)
        self.dim = len(lvals)
        self.n_sites = prod(lvals)
        self.dirs = "xyz"[: self.dim]
        # static background charges (QED_model.py:56-59, 96-101; the bg dof
        # leads every dressed site and each site keeps its own charge's states)
        bg = int(max(abs(q) for q in bg_list)) if bg_list is not None else 0
        self.bg_list = list(bg_list) if bg else None
        basis, states = qed_ops.gauge_invariant_states(spin, pure_theory, self.dim, background=bg)
        ops = qed_ops.dressed_site_operators(spin, pure_theory, self.dim, background=bg)
        self.labels = [
            lat.site_label(lat.zig_zag(self.lvals, i), self.lvals, self.has_obc, not pure_theory)
            for i in range(self.n_sites)
        ]
        cols = None
        if bg:
            cols = [np.flatnonzero(np.isclose(states[l][:, 0], q, rtol=0.0, atol=1e-12))
                    for l, q in zip(self.labels, self.bg_list)]
        self.ops, self.loc_dims = project_operators(ops, basis, self.labels, site_cols=cols)
        self.gauge_states = states
        # symmetry generators (QED_model.py:117-199)
        dg = lambda name: np.real(np.einsum("sii->si", self.ops[name]))
        pairs = lat.link_pairs(self.lvals, self.has_obc)
        link_diags = np.stack([np.stack([dg(f"E_p{d}"), -dg(f"E_m{d}")]) for d in self.dirs])
        link_sectors = np.zeros(self.dim)
        self.pairs, self.link_diags = pairs, link_diags
        if not enumerate_sector:
            return
        if not self.link_symmetries:
            link_diags = None
        if not pure_theory:
            glob = dg("N")[None]
            self.configs = sector_configs(self.loc_dims, glob, np.array([self.n_sites // 2], float),
                                          "U", link_diags, link_sectors, pairs)
        elif not any(self.has_obc) and self.dim == 2:
            lines = lat.origin_lines_2d(self.lvals)
            line_of = {"x": "y", "y": "x"}
            nb_sites = [lines[line_of[d]] for d in self.dirs]
            nb_diags = np.stack([dg(f"E_p{d}") for d in self.dirs])
            self.configs = sector_configs(self.loc_dims, None, None, "U", link_diags, link_sectors,
                                          pairs, nb_diags, np.zeros(self.dim), nb_sites, "U")
        elif not any(self.has_obc) and self.dim == 3:
            faces = lat.origin_faces_3d(self.lvals)
            face_of = {"x": "yz", "y": "xz", "z": "xy"}
            self.configs = sector_configs(self.loc_dims, None, None, "U", link_diags, link_sectors, pairs,
                                          np.stack([dg(f"E_p{d}") for d in self.dirs]), np.zeros(3),
                                          [faces[face_of[d]] for d in self.dirs], "U")
        else:
            self.configs = sector_configs(self.loc_dims, None, None, "U", link_diags,
                                          link_sectors, pairs)
        self.keys = encode(self.configs, self.loc_dims)
        self.N = self.configs.shape[0]

    def label_config(self, name):
        """Local-state indices of the 1D spin-1 benchmark product states named
        by QED_Model.overlap_QMB_state (QED_model.py:796-846), found from their
        physical content in this oracle's own local bases rather than from the
        reference's index table: "V" has every link at E = 0, even sites empty
        and odd sites filled (no net charge anywhere); "meson_center" turns the
        central (even, odd) pair into a meson (even site filled with its right
        link at E = +1, odd site empty with its left link at E = +1)."""
        if self.dim != 1 or self.pure:
            raise NotImplementedError("1D QED with matter only")

        def index(site, n, e_m, e_p):
            st = self.gauge_states[self.labels[site]]
            hit = np.flatnonzero((st[:, 0] == n) & (st[:, 1] == e_m) & (st[:, 2] == e_p))
            assert hit.size == 1
            return int(hit[0])

        cfg = [index(i, i % 2, 0, 0) for i in range(self.n_sites)]
        if name == "meson_center":
            e = self.n_sites // 2
            e -= e % 2
            cfg[e], cfg[e + 1] = index(e, 1, 0, 1), index(e + 1, 0, 1, 0)
        return np.array(cfg)

    def couplings(self, g, m=None, theta=0.0):
        """QED_Hamiltonian_couplings (QED_model.py:719-776)."""
        c = {"g": g, "E": g / 2, "B": -1 / (2 * g), "theta": -theta * g}
        if not self.pure and m is not None:
            t = 0.5
            c.update({"tx_even": -1j * t, "tx_odd": -1j * t, "ty_even": -t, "ty_odd": t,
                      "tz_even": -1j * t, "tz_odd": 1j * t, "m_even": m, "m_odd": -m})
        return c

    def triplets(self, g, m=None, theta=0.0):
        """build_truncated_Hamiltonian term order (QED_model.py:235-394)."""
        c = self.couplings(g, m, theta)
        T = Triplets(self.N)
        args = (self.configs, self.keys, self.loc_dims)
        add_local(T, self.ops["E2"], *args, c["E"])
        if self.dim > 1:
            names = ["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"]
            plaqs = [p for p in (lat.plaquette_sites(lat.zig_zag(self.lvals, i), self.lvals, ["x", "y"],
                                                     self.has_obc) for i in range(self.n_sites)) if p]
            add_nbody_list(T, [self.ops[n] for n in names], plaqs, *args, c["B"], True)
            if self.dim == 3:
                for ax, nm in ((["x", "z"], ["C_px,pz", "C_pz,mx", "C_mz,px", "C_mx,mz"]),
                               (["y", "z"], ["C_py,pz", "C_pz,my", "C_mz,py", "C_my,mz"])):
                    pl = [p for p in (lat.plaquette_sites(lat.zig_zag(self.lvals, i), self.lvals, ax,
                                                          self.has_obc) for i in range(self.n_sites)) if p]
                    add_nbody_list(T, [self.ops[n] for n in nm], pl, *args, c["B"], True)
            if self.dim == 2 and abs(c["theta"]) > 10e-10:
                add_nbody_list(T, [self.ops[n] for n in names], plaqs, *args, c["theta"], True)
            if self.dim == 3 and abs(c["theta"]) > 10e-10:
                for ax, nm in ((["x", "y"], ["EzC_px,py", "C_py,mx", "C_my,px", "C_mx,my"]),
                               (["x", "z"], ["EyC_px,pz", "C_pz,mx", "C_mz,px", "C_mx,mz"]),
                               (["y", "z"], ["ExC_py,pz", "C_pz,my", "C_mz,py", "C_my,mz"])):
                    pl = [p for p in (lat.plaquette_sites(lat.zig_zag(self.lvals, i), self.lvals, ax,
                                                          self.has_obc) for i in range(self.n_sites)) if p]
                    add_nbody_list(T, [self.ops[n] for n in nm], pl, *args, c["theta"], True)
        if not self.pure:
            for stag in ["even", "odd"]:
                mask = lat.staggered_mask(self.lvals, stag)
                ms = [mask[lat.zig_zag(self.lvals, i)] for i in range(self.n_sites)]
                add_local(T, self.ops["N"], *args, c[f"m_{stag}"], ms)
            for d in self.dirs:
                for stag in ["even", "odd"]:
                    mask = lat.staggered_mask(self.lvals, stag)
                    bonds = []
                    for i in range(self.n_sites):
                        co = lat.zig_zag(self.lvals, i)
                        b = lat.neighbor_sites(co, self.lvals, d, self.has_obc)
                        if b is not None and mask[co]:
                            bonds.append(b)
                    if bonds:
                        add_nbody_list(T, [self.ops[f"Q_p{d}_dag"], self.ops[f"Q_m{d}"]], bonds, *args,
                                       c[f"t{d}_{stag}"], True)
        if not self.link_symmetries:
            # 100 E2 - 100 sum_links E_p(i) E_m(i + d) = 50 sum (E_p - E_m)^2 + boundary terms
            add_local(T, self.ops["E2"], *args, 100.0)
            for d in self.dirs:
                bonds = []
                for i in range(self.n_sites):
                    b = lat.neighbor_sites(lat.zig_zag(self.lvals, i), self.lvals, d, self.has_obc)
                    if b is not None:
                        bonds.append(b)
                if bonds:
                    add_nbody_list(T, [self.ops[f"E_p{d}"], self.ops[f"E_m{d}"]], bonds, *args, -100.0, False)
        return T

    def hamiltonian(self, g, m=None, theta=0.0):
        real = self.pure and abs(theta) <= 1e-12
        return self.triplets(g, m, theta).csr(real=real)


def lowest_eigs(H, k, v0=None, tol=0.0):
    """eigsh(which='SA') sorted ascending (qmb_hamiltonian.py:306-317)."""
    if H.shape[0] <= max(2 * k + 2, 64):
        w = np.linalg.eigvalsh(H.toarray())
        return np.sort(w)[:k]
    w = eigsh(H, k=k, which="SA", v0=v0, tol=tol, return_eigenvectors=False)
    return np.sort(w)


def expm_path(H, psi0, t0, t1, num):
    """expm_multiply(-iH, psi0, t0..t1, num, endpoint) (qmb_hamiltonian.py:377-386)."""
    return expm_multiply(-1j * H, psi0, start=t0, stop=t1, num=num, endpoint=True)


class ZnOracle:
    """Pure Z_n gauge theory in 2D (oracle): local basis / operators restated from
    Zn_operators.py:130-360, link constraint E_p(i) + E_m(i+d) = 0, optional
    flux sector sum_{origin line} E_p{d} = f_d (mod n),
    H = -J sum_plaq (C_px,py C_py,mx C_my,px C_mx,my + h.c.) - h sum_sites E_clock."""

    def __init__(self, n, lvals, has_obc, flux_sector=(0, 0)):
        from . import zn_ops
        self.n = n
        self.lvals, self.has_obc = list(lvals), list(has_obc)
        self.n_sites = prod(lvals)
        self.dim = 2
        basis, states = zn_ops.gauge_invariant_states(n)
        ops = zn_ops.dressed_site_operators(n)
        self.labels = [lat.site_label(lat.zig_zag(self.lvals, i), self.lvals, self.has_obc, False)
                       for i in range(self.n_sites)]
        self.ops, self.loc_dims = project_operators(ops, basis, self.labels)
        clock = {l: (basis[l].T @ zn_ops.clock_operator(n, l) @ basis[l]).toarray() for l in set(self.labels)}
        dmax = int(self.loc_dims.max())
        self.ops["E_clock"] = np.zeros((self.n_sites, dmax, dmax), dtype=complex)
        for j, l in enumerate(self.labels):
            d = self.loc_dims[j]
            self.ops["E_clock"][j, :d, :d] = clock[l]
        dg = lambda name: np.real(np.einsum("sii->si", self.ops[name]))
        pairs = lat.link_pairs(self.lvals, self.has_obc)
        link_diags = np.stack([np.stack([dg(f"E_p{d}"), dg(f"E_m{d}")]) for d in "xy"])
        if flux_sector is not None and not any(self.has_obc):
            lines = lat.origin_lines_2d(self.lvals)
            nb_sites = [lines["y"], lines["x"]]
            nb_diags = np.stack([dg("E_px"), dg("E_py")])
            self.configs = sector_configs(self.loc_dims, None, None, "U", link_diags, np.zeros(2), pairs,
                                          nb_diags, np.array(flux_sector, float), nb_sites, f"M{n}")
        else:
            self.configs = sector_configs(self.loc_dims, None, None, "U", link_diags, np.zeros(2), pairs)
        self.keys = encode(self.configs, self.loc_dims) if np.prod(self.loc_dims.astype(float)) < 2**62 else None
        self.N = self.configs.shape[0]

    def hamiltonian(self, J, h):
        if self.keys is None:
            raise NotImplementedError("oracle triplets need 64-bit keys")
        T = Triplets(self.N)
        args = (self.configs, self.keys, self.loc_dims)
        add_local(T, self.ops["E_clock"], *args, -h)
        names = ["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"]
        plaqs = [p for p in (lat.plaquette_sites(lat.zig_zag(self.lvals, i), self.lvals, ["x", "y"],
                                                 self.has_obc) for i in range(self.n_sites)) if p]
        add_nbody_list(T, [self.ops[k] for k in names], plaqs, *args, -J, True)
        return T.csr(real=True)


class SU2Oracle:
    """SU2_Model restated (models/SU2_model.py:31-171,
    200-355, 357-484, 555-624): the hard-coded spin-1/2 operators, or the
    generalised ones (generic=True; any rishon truncation spin), 1D-3D."""

    def __init__(self, pure_theory, lvals, has_obc, sectors=None, spin=0.5, generic=False, bg_list=None):
        from sympy import Rational
        from . import su2_ops
        self.pure = pure_theory
        self.lvals, self.has_obc = list(lvals), list(has_obc)
        self.dim = len(lvals)
        self.n_sites = prod(lvals)
        self.dirs = "xyz"[: self.dim]
        j_max = Rational(int(round(2 * spin)), 2)
        self.generic = bool(generic) or j_max >= 1
        # static background charges (SU2_model.py:63-64; quantum_model.py:440-466)
        bg = Rational(int(round(2 * max(bg_list))), 2) if bg_list is not None else 0
        self.bg_list = list(bg_list) if bg else None
        basis, bgcol = su2_ops.gauge_invariant_states(pure_theory, self.dim, j_max, bg, return_bg=True)
        ops = (su2_ops.dressed_site_operators_gen(pure_theory, self.dim, j_max, bg) if self.generic
               else su2_ops.dressed_site_operators(pure_theory, self.dim, bg))
        self.labels = [lat.site_label(lat.zig_zag(self.lvals, i), self.lvals, self.has_obc, False)
                       for i in range(self.n_sites)]
        self.site_basis = basis
        cols = None
        if bg:
            cols = [np.flatnonzero(np.isclose(bgcol[l], q, rtol=0.0, atol=1e-12))
                    for l, q in zip(self.labels, self.bg_list)]
        self.ops, self.loc_dims = project_operators(ops, basis, self.labels, site_cols=cols)
        dg = lambda name: np.real(np.einsum("sii->si", self.ops[name]))
        pairs = lat.link_pairs(self.lvals, self.has_obc)
        link_diags = np.stack([np.stack([dg(f"T2_p{d}"), -dg(f"T2_m{d}")]) for d in self.dirs])
        link_sectors = np.zeros(self.dim)
        if not pure_theory:
            self.configs = sector_configs(self.loc_dims, dg("N_tot")[None], np.array([float(sectors[0])]),
                                          "U", link_diags, link_sectors, pairs)
        elif not any(self.has_obc) and self.dim == 2:
            lines = lat.origin_lines_2d(self.lvals)
            line_of = {"x": "y", "y": "x"}
            self.configs = sector_configs(self.loc_dims, None, None, "U", link_diags, link_sectors, pairs,
                                          np.stack([dg(f"P_p{d}") for d in self.dirs]), np.ones(self.dim),
                                          [lines[line_of[d]] for d in self.dirs], "Z")
        elif not any(self.has_obc) and self.dim == 3:
            faces = lat.origin_faces_3d(self.lvals)
            face_of = {"x": "yz", "y": "xz", "z": "xy"}
            self.configs = sector_configs(self.loc_dims, None, None, "U", link_diags, link_sectors, pairs,
                                          np.stack([dg(f"P_p{d}") for d in self.dirs]), np.ones(3),
                                          [faces[face_of[d]] for d in self.dirs], "Z")
        else:
            self.configs = sector_configs(self.loc_dims, None, None, "U", link_diags, link_sectors, pairs)
        self.keys = encode(self.configs, self.loc_dims)
        self.N = self.configs.shape[0]

    def couplings(self, g, m=None, theta=0.0):
        """SU2_Hamiltonian_couplings (SU2_model.py:555-624)."""
        c = {"E": 8 * g / 3 if self.dim == 1 else g / 2, "B": 0 if self.dim == 1 else -1 / (2 * g),
             "theta": -theta * g}
        if not self.pure and m is not None:
            t = 2 * np.sqrt(2)
            c.update({"tx_even": -1j * t, "tx_odd": -1j * t, "ty_even": -t, "ty_odd": t,
                      "tz_even": -t, "tz_odd": t, "m_even": m, "m_odd": -m})
        return c

    def _plaqs(self, axes):
        return [p for p in (lat.plaquette_sites(lat.zig_zag(self.lvals, i), self.lvals, list(axes), self.has_obc)
                            for i in range(self.n_sites)) if p]

    def _bonds(self, d, stag):
        mask = lat.staggered_mask(self.lvals, stag)
        bonds = []
        for i in range(self.n_sites):
            co = lat.zig_zag(self.lvals, i)
            b = lat.neighbor_sites(co, self.lvals, d, self.has_obc)
            if b is not None and mask[co]:
                bonds.append(b)
        return bonds

    def triplets(self, g, m=None, theta=0.0, noise=0.0):
        """build_base_Hamiltonian (SU2_model.py:200-355) or build_gen_Hamiltonian
        (:357-484) term order."""
        if self.generic:
            return self._triplets_gen(g, m)
        c = self.couplings(g, m, theta)
        T = Triplets(self.N)
        args = (self.configs, self.keys, self.loc_dims)
        add_local(T, self.ops["E_square"], *args, c["E"])
        planes = [] if self.dim == 1 else ["xy"] if self.dim == 2 else ["xy", "xz", "yz"]
        for a, b in planes:
            names = [f"C_p{a},p{b}", f"C_p{b},m{a}", f"C_m{b},p{a}", f"C_m{a},m{b}"]
            add_nbody_list(T, [self.ops[n] for n in names], self._plaqs((a, b)), *args, c["B"], True)
        if not self.pure:
            for stag in ["even", "odd"]:
                mask = lat.staggered_mask(self.lvals, stag)
                ms = [mask[lat.zig_zag(self.lvals, i)] for i in range(self.n_sites)]
                add_local(T, self.ops["N-1"], *args, c[f"m_{stag}"], ms)
            for d in self.dirs:
                for stag in ["even", "odd"]:
                    mask = lat.staggered_mask(self.lvals, stag)
                    bonds = []
                    for i in range(self.n_sites):
                        co = lat.zig_zag(self.lvals, i)
                        b = lat.neighbor_sites(co, self.lvals, d, self.has_obc)
                        if b is not None and mask[co]:
                            bonds.append(b)
                    if bonds:
                        add_nbody_list(T, [self.ops[f"Qp{d}_dag"], self.ops[f"Qm{d}"]], bonds, *args,
                                       c[f"t{d}_{stag}"], True)
        if self.dim == 3 and abs(c["theta"]) > 10e-10:
            for names, axes in ((["EzC_px,py", "C_py,mx", "C_my,px", "C_mx,my"], "xy"),
                                (["EyC_px,pz", "C_pz,mx", "C_mz,px", "C_mx,mz"], "xz"),
                                (["ExC_py,pz", "C_pz,my", "C_mz,py", "C_my,mz"], "yz")):
                add_nbody_list(T, [self.ops[n] for n in names], self._plaqs(axes), *args, c["theta"], True)
        if self.bg_list is not None and noise > 0 and not any(self.has_obc):
            bonds = [b for b in (lat.neighbor_sites(lat.zig_zag(self.lvals, i), self.lvals, "x", self.has_obc)
                                 for i in range(self.n_sites)) if b is not None]
            add_nbody_list(T, [self.ops["Vpx_dag"], self.ops["Vmx"]], bonds, *args, -noise, True)
        return T

    PLAQ_SETS = [["AB", "AB", "AB", "AB"], ["AA", "AB", "BB", "AB"], ["AB", "AB", "AA", "BB"],
                 ["AA", "AB", "BA", "BB"], ["AB", "BB", "AB", "AA"], ["AA", "BB", "BB", "AA"],
                 ["AB", "BB", "AA", "BA"], ["AA", "BB", "BA", "BA"], ["BB", "AA", "AB", "AB"],
                 ["BA", "AA", "BB", "AB"], ["BB", "AA", "AA", "BB"], ["BA", "AA", "BA", "BB"],
                 ["BB", "BA", "AB", "AA"], ["BA", "BA", "BB", "AA"], ["BB", "BA", "AA", "BA"],
                 ["BA", "BA", "BA", "BA"]]

    def _triplets_gen(self, g, m=None):
        c = self.couplings(g, m)
        T = Triplets(self.N)
        args = (self.configs, self.keys, self.loc_dims)
        add_local(T, self.ops["E_square"], *args, c["E"])
        if not self.pure:
            for stag in ["even", "odd"]:
                mask = lat.staggered_mask(self.lvals, stag)
                ms = [mask[lat.zig_zag(self.lvals, i)] for i in range(self.n_sites)]
                add_local(T, self.ops["N_tot"], *args, c[f"m_{stag}"], ms)
            for d in self.dirs:
                for stag in ["even", "odd"]:
                    bonds = self._bonds(d, stag)
                    if not bonds:
                        continue
                    for names in ([f"Q1_p{d}_dag", f"Q2_m{d}"], [f"Q2_p{d}_dag", f"Q1_m{d}"]):
                        add_nbody_list(T, [self.ops[n] for n in names], bonds, *args, c[f"t{d}_{stag}"], True)
        planes = [] if self.dim == 1 else ["xy"] if self.dim == 2 else ["xy", "xz", "yz"]
        for a, b in planes:
            plaqs = self._plaqs((a, b))
            for ps in self.PLAQ_SETS:
                names = [f"C{ps[0]}_p{a},p{b}", f"C{ps[1]}_p{b},m{a}", f"C{ps[2]}_m{b},p{a}",
                         f"C{ps[3]}_m{a},m{b}"]
                add_nbody_list(T, [self.ops[n] for n in names], plaqs, *args, c["B"], True)
        return T

    def hamiltonian(self, g, m=None, theta=0.0, noise=0.0):
        return self.triplets(g, m, theta, noise).csr(real=self.pure and abs(theta) <= 1e-12)
