"""SU(2) spin-1/2 dressed sites (oracle; test infrastructure).

Restates, with SymPy's Clebsch-Gordan coefficients as the reference uses them:
  * couple_two_spins / add_new_spin / group_sorted_spin_configs
                                   operators/SU2_singlets.py:30-60, 172-300
  * get_spin_Hilbert_spaces / SU2_singlet_canonical_vector
                                   operators/SU2_singlets.py:303-420
  * SU2_gauge_invariant_states     operators/SU2_operators.py:625-714
  * SU2_Rishon (j = 1/2)           operators/SU2_rishons.py:50-80
  * SU2_generators                 operators/spin_operators.py:91-127
  * fermi_operators (two colours)  operators/bose_fermi_operators.py:14-70
  * SU2_dressed_site_operators     operators/SU2_operators.py:35-300 (1D, 2D)
Kronecker order follows qmb_operator (modeling/qmb_operations.py:28-73):
the first listed factor is the most significant.
"""
from itertools import product

import numpy as np
import scipy.sparse as sp
from sympy import S
from sympy.physics.wigner import clebsch_gordan

from . import lattice as lat


def _mvals(j):
    return [j - k for k in range(int(2 * j) + 1)]


def _couple(j1, j2, singlet, M=None):
    """couple_two_spins (SU2_singlets.py:30-60)."""
    m1s = [M] if M is not None else _mvals(j1)
    j3s = [S(0)] if singlet else [abs(j1 - j2) + k for k in range(int(j1 + j2 - abs(j1 - j2)) + 1)]
    out = []
    for m1, m2, j3 in product(m1s, _mvals(j2), j3s):
        m3 = m1 + m2
        if abs(m3) > j3:
            continue
        cg = clebsch_gordan(S(j1), S(j2), S(j3), S(m1), S(m2), S(m3))
        if cg != 0:
            out.append([S(j2), S(m2), cg, S(j3), S(m3)] if M is not None
                       else [S(j1), S(m1), S(j2), S(m2), cg, S(j3), S(m3)])
    return out


def singlets(spins):
    """get_SU2_singlets: list of (J tuple, [(m tuple, cg product)])."""
    cfgs = _couple(spins[0], spins[1], len(spins) == 2)
    for k in range(2, len(spins)):
        cfgs = [c + n for c in cfgs for n in _couple(c[-2], spins[k], k == len(spins) - 1, M=c[-1])]
    if not cfgs:
        return []
    L = len(cfgs[0])
    m_idx = [1] + list(range(3, L, 5))
    cg_idx = list(range(4, L, 5))
    j_idx = list(range(5, L, 5))
    cfgs = sorted(cfgs, key=lambda c: tuple(c[k] for k in j_idx))
    groups = []
    for c in cfgs:
        J = tuple(c[k] for k in j_idx)
        amp = S(1)
        for k in cg_idx:
            amp *= c[k]
        ms = tuple(c[k] for k in m_idx)
        if groups and groups[-1][0] == J:
            groups[-1][1].append((ms, amp))
        else:
            groups.append((J, [(ms, amp)]))
    return groups


def gauge_invariant_states(pure_theory, dim):
    """SU2_gauge_invariant_states for j_max = 1/2: basis[label] (csr, D x d)."""
    n_link = 2 * dim
    links = [f"m{a}" for a in "xyz"[:dim]] + [f"p{a}" for a in "xyz"[:dim]]
    rish_j, rish_m = [S(0), S(1) / 2, S(1) / 2], [S(0), S(1) / 2, -S(1) / 2]
    j_list = [rish_j] * n_link
    m_list = [rish_m] * n_link
    if not pure_theory:
        j_list = [["V", S(1) / 2, S(1) / 2, "P"]] + j_list
        m_list = [[S(0), S(1) / 2, -S(1) / 2, S(0)]] + m_list
    full = list(zip(product(*j_list), product(*m_list)))
    spins = [[S(0), S(1) / 2]] * n_link
    if not pure_theory:
        spins = [[S(0), S(1) / 2, S(0)]] + spins
    labels = ["site"] + [f"site_{b}" for b in lat.border_labels(dim)]
    cols = {l: [] for l in labels}
    n_after = int(np.prod([len(s) for s in spins[1:]]))
    for ii, cfg in enumerate(product(*spins)):
        cfg = list(cfg)
        j2 = [int(2 * x) for x in cfg]
        tot = sum(j2)
        if tot % 2 or max(j2) > tot - max(j2):
            continue
        Jcfg = list(cfg)
        if not pure_theory:
            sector = (ii // n_after) % 3
            if sector == 0:
                Jcfg[0] = "V"
            elif sector == 2:
                Jcfg[0] = "P"
        for _, terms in singlets(cfg):
            vec = np.zeros(len(full))
            jm = {tuple(Jcfg) + ms: amp for ms, amp in terms}
            for k, (jc, mc) in enumerate(full):
                key = tuple(jc) + tuple(mc)
                if key in jm:
                    vec[k] = float(jm[key])
            link_j = cfg[1:] if not pure_theory else cfg
            zero = {links[k] for k in range(n_link) if link_j[k] == 0}
            for l in labels:
                parts = l.split("_")[1:]
                if all(p in zero for p in parts):
                    cols[l].append(vec)
    return {l: sp.csr_matrix(np.column_stack(c)) for l, c in cols.items() if c}


def _rishon():
    cf = 1 / 2 ** 0.25
    P = sp.diags([1.0, -1.0, -1.0]).tocsr()
    Zr = sp.csr_matrix(([cf, cf], ([0, 2], [1, 0])), shape=(3, 3))
    Zg = sp.csr_matrix(([cf, -cf], ([0, 1], [2, 0])), shape=(3, 3))
    ops = {"P": P, "Iz": sp.identity(3, format="csr"), "Zr": Zr, "Zg": Zg}
    for c in "rg":
        Z = ops[f"Z{c}"]
        ops[f"Z{c}_dag"] = Z.T.tocsr()
        ops[f"Z{c}_P"] = Z @ P
        ops[f"P_Z{c}_dag"] = P @ Z.T
        ops[f"Z{c}_dag_P"] = Z.T @ P
    # SU2_generators(1/2): block_diag(0, spin-1/2 matrices)
    Tz = sp.block_diag([sp.csr_matrix((1, 1)), sp.diags([0.5, -0.5])]).tocsr()
    Tp = sp.block_diag([sp.csr_matrix((1, 1)), sp.csr_matrix(np.array([[0.0, 1.0], [0.0, 0.0]]))]).tocsr()
    ops["Tz"], ops["Tp"], ops["Tm"] = Tz, Tp, Tp.T.tocsr()
    ops["Tx"] = 0.5 * (Tp + Tp.T)
    ops["Ty"] = -0.5j * (Tp - Tp.T)
    ops["T2"] = sp.block_diag([sp.csr_matrix((1, 1)), sp.diags([0.75, 0.75])]).tocsr()
    return ops


def _matter():
    psi = sp.csr_matrix(np.array([[0.0, 1.0], [0.0, 0.0]]))
    Pp = sp.diags([1.0, -1.0]).tocsr()
    I2 = sp.identity(2, format="csr")
    N = psi.T @ psi
    ops = {}
    ops["psi_g"] = sp.kron(psi, I2).tocsr()      # psi_down
    ops["psi_r"] = sp.kron(Pp, psi).tocsr()      # psi_up
    ops["N_g"] = sp.kron(N, I2).tocsr()
    ops["N_r"] = sp.kron(I2, N).tocsr()
    ops["N_pair"] = ops["N_r"] @ ops["N_g"]
    ops["N_tot"] = ops["N_r"] + ops["N_g"]
    ops["N_single"] = ops["N_tot"] - 2 * ops["N_pair"]
    ops["N_zero"] = sp.identity(4) - ops["N_pair"] - ops["N_single"]
    ops["ID_psi"] = sp.identity(4, format="csr")
    ops["P_psi"] = sp.kron(Pp, Pp).tocsr()
    for c in "rg":
        ops[f"psi_{c}_dag"] = ops[f"psi_{c}"].T.tocsr()
        ops[f"psi_{c}_dag_P"] = ops[f"psi_{c}_dag"] @ ops["P_psi"]
    return ops


def _kron(ops, names):
    out = ops[names[0]]
    for n in names[1:]:
        out = sp.kron(out, ops[n])
    return sp.csr_matrix(out)


def dressed_site_operators(pure_theory, dim):
    """SU2_dressed_site_operators(1/2, pure_theory, dim) for dim 1, 2."""
    ins = _rishon()
    if not pure_theory:
        ins.update(_matter())
    n_link = 2 * dim
    ops = {}
    for op in ["T2", "P", "Tx", "Ty", "Tz"]:
        for k, l in enumerate([f"m{a}" for a in "xyz"[:dim]] + [f"p{a}" for a in "xyz"[:dim]]):
            names = ["Iz"] * n_link
            names[k] = op
            ops[f"{op}_{l}"] = _kron(ins, names)
    if dim == 1:
        ops["U"] = sum(_kron(ins, [f"Z{c}_dag_P", f"Z{c}"]) for c in "rg") / np.sqrt(2)
    elif dim == 2:
        ops["C_px,py"] = sum(_kron(ins, ["Iz", "Iz", f"Z{c}_P", f"Z{c}_dag"]) for c in "rg")
        ops["C_py,mx"] = sum(_kron(ins, [f"P_Z{c}_dag", "P", "P", f"Z{c}"]) for c in "rg")
        ops["C_mx,my"] = sum(_kron(ins, [f"Z{c}_P", f"Z{c}_dag", "Iz", "Iz"]) for c in "rg")
        ops["C_my,px"] = sum(_kron(ins, ["Iz", f"Z{c}_P", f"Z{c}_dag", "Iz"]) for c in "rg")
    else:
        raise NotImplementedError
    if not pure_theory:
        for k in list(ops):
            ops[k] = sp.kron(ins["ID_psi"], ops[k]).tocsr()
        if dim == 1:
            lists = {"Qpx_dag": lambda c: [f"psi_{c}_dag_P", "P", f"Z{c}"],
                     "Qmx_dag": lambda c: [f"psi_{c}_dag_P", f"Z{c}", "Iz"]}
        else:
            lists = {"Qmx_dag": lambda c: [f"psi_{c}_dag_P", f"Z{c}", "Iz", "Iz", "Iz"],
                     "Qmy_dag": lambda c: [f"psi_{c}_dag_P", "P", f"Z{c}", "Iz", "Iz"],
                     "Qpx_dag": lambda c: [f"psi_{c}_dag_P", "P", "P", f"Z{c}", "Iz"],
                     "Qpy_dag": lambda c: [f"psi_{c}_dag_P", "P", "P", "P", f"Z{c}"]}
        for name, fn in lists.items():
            ops[name] = sum(_kron(ins, fn(c)) for c in "rg")
        for name in list(lists):
            ops[name.replace("_dag", "")] = ops[name].conj().T.tocsr()
        for lab in ["r", "g", "tot", "single", "pair", "zero"]:
            ops[f"N_{lab}"] = _kron(ins, [f"N_{lab}"] + ["Iz"] * n_link)
        ops["N-1"] = ops["N_tot"] - sp.identity(ops["N_tot"].shape[0])
    ops["E_square"] = sum(0.5 * ops[k] for k in ops if k.startswith("T2_"))
    return ops


def total_spin_operators(pure_theory, dim):
    """Sum of the site's SU(2) generators (matter + every half-link), the
    Gauss-law generators whose common kernel is the singlet space."""
    ins = _rishon()
    n_link = 2 * dim
    out = {}
    for a in ("x", "y", "z"):
        tot = 0
        for k in range(n_link):
            names = ["Iz"] * n_link
            names[k] = f"T{a}"
            term = _kron(ins, names)
            if not pure_theory:
                term = sp.kron(sp.identity(4), term)
            tot = tot + term
        if not pure_theory:
            # matter spin: the single-fermion states carry spin 1/2 (r = up, g = down)
            sz = sp.diags([0.0, 0.5, -0.5, 0.0])
            sp_ = sp.csr_matrix(([1.0], ([1], [2])), shape=(4, 4))
            mat = {"z": sz, "x": 0.5 * (sp_ + sp_.T), "y": -0.5j * (sp_ - sp_.T)}[a]
            tot = tot + sp.kron(mat, sp.identity(3 ** n_link))
        out[a] = sp.csr_matrix(tot)
    return out
