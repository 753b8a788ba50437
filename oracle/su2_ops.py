"""SU(2) dressed sites (oracle; test infrastructure).

Restates, with SymPy's Clebsch-Gordan coefficients as the reference uses them:
  * couple_two_spins / add_new_spin / group_sorted_spin_configs
                                   operators/SU2_singlets.py:30-60, 172-300
  * get_spin_Hilbert_spaces / SU2_singlet_canonical_vector
                                   operators/SU2_singlets.py:303-420
  * SU2_gauge_invariant_states     operators/SU2_operators.py:625-714
  * SU2_Rishon (j = 1/2)           operators/SU2_rishons.py:50-80
  * SU2_Rishon_gen (any j_max)     operators/SU2_rishons.py:106-234
  * SU2_generators                 operators/spin_operators.py:75-127
  * fermi_operators (two colours)  operators/bose_fermi_operators.py:14-70
  * SU2_dressed_site_operators     operators/SU2_operators.py:35-300 (1D-3D)
  * SU2_gen_dressed_site_operators operators/SU2_operators.py:302-592 (1D-3D)
Pins (tests): the rishon algebra the reference checks (SU2_rishons.py:
80-103, 179-194), singlet spaces equal to the kernel of the total SU(2)
generators, and the generalised and hard-coded spin-1/2 models sharing their
sector and spectrum (validation/su2/test_su2_hamiltonians.py).
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


def _irreps(j_max):
    return [S(k) / 2 for k in range(int(2 * j_max) + 1)]


def gauge_invariant_states(pure_theory, dim, j_max=S(1) / 2, background=0, return_bg=False):
    """SU2_gauge_invariant_states: basis[label] (csr, D x d); with a
    background (spin irreps 0..background as the leading, first-coupled dof)
    and return_bg, also the background spin of every column per label."""
    n_link = 2 * dim
    links = [f"m{a}" for a in "xyz"[:dim]] + [f"p{a}" for a in "xyz"[:dim]]
    rish_j = [j for j in _irreps(j_max) for _ in _mvals(j)]
    rish_m = [m for j in _irreps(j_max) for m in _mvals(j)]
    j_list = [rish_j] * n_link
    m_list = [rish_m] * n_link
    if not pure_theory:
        j_list = [["V", S(1) / 2, S(1) / 2, "P"]] + j_list
        m_list = [[S(0), S(1) / 2, -S(1) / 2, S(0)]] + m_list
    bg = S(background)
    if bg:
        j_list = [[j for j in _irreps(bg) for _ in _mvals(j)]] + j_list
        m_list = [[m for j in _irreps(bg) for m in _mvals(j)]] + m_list
    full = {tuple(jc) + tuple(mc): k for k, (jc, mc) in enumerate(zip(product(*j_list), product(*m_list)))}
    spins = [_irreps(j_max)] * n_link
    if not pure_theory:
        spins = [[S(0), S(1) / 2, S(0)]] + spins
    vind = 0
    if bg:
        spins = [_irreps(bg)] + spins
        vind = 1
    labels = ["site"] + [f"site_{b}" for b in lat.border_labels(dim)]
    cols = {l: [] for l in labels}
    bgs = {l: [] for l in labels}
    n_after = int(np.prod([len(s) for s in spins[vind + 1:]]))
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
                Jcfg[vind] = "V"
            elif sector == 2:
                Jcfg[vind] = "P"
        for _, terms in singlets(cfg):
            vec = np.zeros(len(full))
            for ms, amp in terms:
                vec[full[tuple(Jcfg) + tuple(ms)]] = float(amp)
            link_j = cfg[vind + (0 if pure_theory else 1):]
            zero = {links[k] for k in range(n_link) if link_j[k] == 0}
            for l in labels:
                parts = l.split("_")[1:]
                if all(p in zero for p in parts):
                    cols[l].append(vec)
                    bgs[l].append(float(cfg[0]) if bg else 0.0)
    out = {l: sp.csr_matrix(np.column_stack(c)) for l, c in cols.items() if c}
    if return_bg:
        return out, {l: np.array(bgs[l]) for l in out}
    return out


def generators(j_max):
    """SU2_generators(j_max): block_diag over the irreps 0..j_max of the spin
    matrices of get_spin_operators (spin_operators.py:31-60, 75-127)."""
    blocks = {"Tz": [sp.csr_matrix((1, 1))], "Tp": [sp.csr_matrix((1, 1))], "T2": [sp.csr_matrix((1, 1))]}
    for j in _irreps(j_max)[1:]:
        jj = float(j)
        mz = np.array([float(m) for m in _mvals(j)])
        blocks["Tz"].append(sp.diags(mz))
        blocks["Tp"].append(sp.diags(np.sqrt(jj * (jj + 1) - mz[1:] * (mz[1:] + 1)), 1,
                                     shape=(len(mz), len(mz))))
        blocks["T2"].append(sp.diags([jj * (jj + 1)] * len(mz)))
    out = {k: sp.block_diag(v, format="csr") for k, v in blocks.items()}
    out["Tm"] = out["Tp"].T.tocsr()
    out["Tx"] = 0.5 * (out["Tp"] + out["Tm"])
    out["Ty"] = -0.5j * (out["Tp"] - out["Tm"])
    return out


def rishon_gen(j_max):
    """SU2_Rishon_gen(j_max).ops: superdiagonal bands of chi coefficients."""
    largest = int(2 * j_max + 1)
    size = sum(range(1, largest + 1))
    shape = (size, size)
    P_diag = []
    for s_size in range(largest):
        P_diag += [(-1) ** s_size] * (s_size + 1)
    ops = {"P": sp.diags(P_diag, 0, shape, format="csr"), "Iz": sp.identity(size, format="csr")}

    def chi(s, color, m):
        num = (s + m + 1) if color == "r" else (s - m + 1)
        return np.sqrt(num / np.sqrt((2 * s + 1) * (2 * s + 2)))

    for color, d0 in (("r", 1), ("g", 2)):
        entries, offsets = [], []
        in_zeros, diag = 0, d0
        for s_size in range(largest - 1):
            spin = s_size / 2
            c = [chi(spin, color, float(m)) for m in _mvals(S(s_size) / 2)]
            out_zeros = size - len(c) - diag - in_zeros
            entries.append([0.0] * in_zeros + c + [0.0] * out_zeros)
            offsets.append(diag)
            diag += 1
            in_zeros += s_size + 1
        Z = sp.diags(entries, offsets, shape, format="csr")
        ops[f"Z{color}"] = Z
        ops[f"Z{color}_dag"] = Z.T.tocsr()
        ops[f"ZB_{color}"] = Z
        ops[f"ZB_{color}_dag"] = Z.T.tocsr()
        ops[f"ZB_{color}_P"] = Z @ ops["P"]
        ops[f"P_ZB_{color}_dag"] = ops["P"] @ Z.T
    ops["ZA_r"] = ops["ZB_g_dag"]
    ops["ZA_g"] = -ops["ZB_r_dag"]
    for color in "rg":
        ops[f"ZA_{color}_dag"] = ops[f"ZA_{color}"].T.tocsr()
        ops[f"ZA_{color}_P"] = ops[f"ZA_{color}"] @ ops["P"]
        ops[f"P_ZA_{color}_dag"] = ops["P"] @ ops[f"ZA_{color}_dag"]
    ops.update(generators(j_max))
    return ops


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


def _background_slot(ops, ins, pure_theory, dim, background):
    """kron(identity(bg), op) for every operator and the background Casimir
    'bg' = T2_bg x identity (SU2_operators.py:270-281, 580-590); ins holds
    the single-dof operators of the set ('Iz' of its rishons)."""
    gen = generators(S(background))
    nbg = gen["T2"].shape[0]
    out = {k: sp.kron(sp.identity(nbg), v).tocsr() for k, v in ops.items()}
    ins = dict(ins)
    ins["T2_bg"] = gen["T2"]
    n_link = 2 * dim
    ids = ["Iz"] * n_link if pure_theory else ["ID_psi"] + ["Iz"] * n_link
    if not pure_theory:
        ins.update(_matter())
    out["bg"] = _kron(ins, ["T2_bg"] + ids)
    return out, ins


def _with_background(ops, pure_theory, dim, background):
    """The background slot and, in 1D with matter, the V hoppings
    (SU2_operators.py:270-298)."""
    out, ins = _background_slot(ops, _rishon(), pure_theory, dim, background)
    if dim == 1 and not pure_theory:
        out["Vmx_dag"] = sum(_kron(ins, [f"Z{c}_dag", "P_psi", f"Z{c}", "Iz"]) for c in "rg")
        out["Vpx_dag"] = sum(_kron(ins, [f"Z{c}_dag", "P_psi", "P", f"Z{c}"]) for c in "rg")
        out["Vmx"] = out["Vmx_dag"].conj().T.tocsr()
        out["Vpx"] = out["Vpx_dag"].conj().T.tocsr()
    return out


def dressed_site_operators(pure_theory, dim, background=0):
    """SU2_dressed_site_operators(1/2, pure_theory, dim, background)."""
    if background:
        return _with_background(dressed_site_operators(pure_theory, dim), pure_theory, dim, background)
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
    else:
        for corner, lst in _corner_lists(dim).items():
            ops[f"C_{corner}"] = sum(_kron(ins, [f.format(a=f"Z{c}", b=f"Z{c}") for f in lst]) for c in "rg")
    if dim == 3:
        # theta-term corners (SU2_operators.py:231-250)
        sig = {"x": 0.5 * np.array([[0, 1], [1, 0]]), "y": 0.5 * np.array([[0, -1j], [1j, 0]]),
               "z": 0.5 * np.array([[1, 0], [0, -1]])}
        for k in ("EzC_px,py", "ExC_py,pz", "EyC_px,pz"):
            ops[k] = 0
        for nu in "xyz":
            Ez = ops[f"T{nu}_pz"] + ops[f"T{nu}_mz"]
            Ey = ops[f"T{nu}_py"] + ops[f"T{nu}_my"]
            Ex = ops[f"T{nu}_px"] + ops[f"T{nu}_mx"]
            for ii, c1 in enumerate("rg"):
                for jj, c2 in enumerate("rg"):
                    factor = sig[nu][jj, ii]
                    ops["EzC_px,py"] = ops["EzC_px,py"] + 1j * _kron(ins, ["Iz", "Iz", "Iz", f"Z{c1}_P", f"Z{c2}_dag", "Iz"]) @ (factor * Ez)
                    ops["EyC_px,pz"] = ops["EyC_px,pz"] - 1j * _kron(ins, ["Iz", "Iz", "Iz", f"Z{c1}_P", "P", f"Z{c2}_dag"]) @ (factor * Ey)
                    ops["ExC_py,pz"] = ops["ExC_py,pz"] + 1j * _kron(ins, ["Iz", "Iz", "Iz", "Iz", f"Z{c1}_P", f"Z{c2}_dag"]) @ (factor * Ex)
    if not pure_theory:
        for k in list(ops):
            ops[k] = sp.kron(ins["ID_psi"], ops[k]).tocsr()
        halves = [f"m{a}" for a in "xyz"[:dim]] + [f"p{a}" for a in "xyz"[:dim]]
        lists = {f"Q{h}_dag": (lambda c, k=k: [f"psi_{c}_dag_P"] + ["P"] * k + [f"Z{c}"] + ["Iz"] * (n_link - k - 1))
                 for k, h in enumerate(halves)}
        for name, fn in lists.items():
            ops[name] = sum(_kron(ins, fn(c)) for c in "rg")
        for name in list(lists):
            ops[name.replace("_dag", "")] = ops[name].conj().T.tocsr()
        for lab in ["r", "g", "tot", "single", "pair", "zero"]:
            ops[f"N_{lab}"] = _kron(ins, [f"N_{lab}"] + ["Iz"] * n_link)
        ops["N-1"] = ops["N_tot"] - sp.identity(ops["N_tot"].shape[0])
    ops["E_square"] = sum(0.5 * ops[k] for k in ops if k.startswith("T2_"))
    return ops


def _corner_lists(dim):
    """Factor lists of the corner operators, half-links in the order m_x,
    m_y, (m_z,) p_x, p_y, (p_z) (SU2_operators.py:104-131, 200-257, 383-399,
    431-492); {a} / {b} stand for the first / second rishon of the corner."""
    if dim == 2:
        return {"px,py": ["Iz", "Iz", "{a}_P", "{b}_dag"], "py,mx": ["P_{b}_dag", "P", "P", "{a}"],
                "mx,my": ["{a}_P", "{b}_dag", "Iz", "Iz"], "my,px": ["Iz", "{a}_P", "{b}_dag", "Iz"]}
    if dim == 3:
        return {"px,py": ["Iz", "Iz", "Iz", "{a}_P", "{b}_dag", "Iz"],
                "py,mx": ["P_{b}_dag", "P", "P", "P", "{a}", "Iz"],
                "mx,my": ["{a}_P", "{b}_dag", "Iz", "Iz", "Iz", "Iz"],
                "my,px": ["Iz", "{a}_P", "P", "{b}_dag", "Iz", "Iz"],
                "px,pz": ["Iz", "Iz", "Iz", "{a}_P", "P", "{b}_dag"],
                "pz,mx": ["P_{b}_dag", "P", "P", "P", "P", "{a}"],
                "mx,mz": ["{a}_P", "P", "{b}_dag", "Iz", "Iz", "Iz"],
                "mz,px": ["Iz", "Iz", "{a}_P", "{b}_dag", "Iz", "Iz"],
                "py,pz": ["Iz", "Iz", "Iz", "Iz", "{a}_P", "{b}_dag"],
                "pz,my": ["Iz", "P_{b}_dag", "P", "P", "P", "{a}"],
                "my,mz": ["Iz", "{a}_P", "{b}_dag", "Iz", "Iz", "Iz"],
                "mz,py": ["Iz", "Iz", "{a}_P", "P", "{b}_dag", "Iz"]}
    return {}


def dressed_site_operators_gen(pure_theory, dim, j_max, background=0):
    """SU2_gen_dressed_site_operators(j_max, pure_theory, dim, background)
    (SU2_operators.py:302-590); a background adds its slot and Casimir only
    (no V hoppings in this set)."""
    if background:
        out, _ = _background_slot(dressed_site_operators_gen(pure_theory, dim, j_max), rishon_gen(j_max),
                                  pure_theory, dim, background)
        return out
    ins = rishon_gen(j_max)
    if not pure_theory:
        ins.update(_matter())
    n_link = 2 * dim
    halves = [f"m{a}" for a in "xyz"[:dim]] + [f"p{a}" for a in "xyz"[:dim]]
    ops = {}
    for op in ["T2", "P"]:
        for k, l in enumerate(halves):
            names = ["Iz"] * n_link
            names[k] = op
            ops[f"{op}_{l}"] = _kron(ins, names)
    for l1, l2 in product("AB", repeat=2):
        for corner, lst in _corner_lists(dim).items():
            ops[f"C{l1}{l2}_{corner}"] = sum(
                _kron(ins, [f.format(a=f"Z{l1}_{c}", b=f"Z{l2}_{c}") for f in lst]) for c in "rg")
    if not pure_theory:
        for k in list(ops):
            ops[k] = sp.kron(ins["ID_psi"], ops[k]).tocsr()
        for k, h in enumerate(halves):
            pre, post = ["P"] * k, ["Iz"] * (n_link - k - 1)
            ops[f"Q1_{h}_dag"] = (_kron(ins, ["psi_r_dag_P"] + pre + ["Zg_dag"] + post)
                                  - _kron(ins, ["psi_g_dag_P"] + pre + ["Zr_dag"] + post))
            ops[f"Q2_{h}_dag"] = (_kron(ins, ["psi_r_dag_P"] + pre + ["Zr"] + post)
                                  + _kron(ins, ["psi_g_dag_P"] + pre + ["Zg"] + post))
        for h in halves:
            for q in ("Q1", "Q2"):
                ops[f"{q}_{h}"] = ops[f"{q}_{h}_dag"].conj().T.tocsr()
        for lab in ["r", "g", "tot", "single", "pair", "zero"]:
            ops[f"N_{lab}"] = _kron(ins, [f"N_{lab}"] + ["Iz"] * n_link)
    ops["E_square"] = sum(0.5 * ops[f"T2_{s}{d}"] for s in "pm" for d in "xyz"[:dim])
    return ops


def total_spin_operators(pure_theory, dim, j_max=S(1) / 2):
    """Sum of the site's SU(2) generators (matter + every half-link), the
    Gauss-law generators whose common kernel is the singlet space."""
    ins = generators(j_max)
    ins["Iz"] = sp.identity(ins["Tz"].shape[0], format="csr")
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
            tot = tot + sp.kron(mat, sp.identity(ins["Iz"].shape[0] ** n_link))
        out[a] = sp.csr_matrix(tot)
    return out
