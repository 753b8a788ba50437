"""Z_n pure-gauge local operators and gauge-invariant bases (oracle; test infrastructure).

Restates edlgt/operators/Zn_operators.py:130-360 (Zn_rishon_operators,
Zn_dressed_site_operators, Zn_gauge_invariant_states, Zn_border_configs) for
the pure theory with scipy Kronecker products in the reference's factor order
(m_x, m_y, p_x, p_y), then adds the clock electric operator of the Z_n
Hamiltonian (derived, the reference has no Z_n model class).
"""
from itertools import product

import numpy as np
from scipy.sparse import csr_matrix, diags, identity, kron


def rishon_operators(n):
    """Zn_rishon_operators(n, pure_theory=True) (Zn_operators.py:130-201)."""
    ops = {}
    ops["U"] = csr_matrix(diags([np.ones(n - 1), np.ones(1)], [-1, n - 1], (n, n)))
    ops["IDz"] = csr_matrix(identity(n))
    ops["n"] = csr_matrix(diags(np.arange(n, dtype=float), 0, (n, n)))
    ops["E"] = csr_matrix(ops["n"] - 0.5 * (n - 1) * identity(n))
    ops["E_square"] = csr_matrix(ops["E"] @ ops["E"])
    ops["P"] = csr_matrix(identity(n))
    ops["Zp"] = ops["U"]
    ops["Zm"] = csr_matrix(ops["P"] @ ops["U"])
    for s in "pm":
        ops[f"Z{s}_dag"] = csr_matrix(ops[f"Z{s}"].transpose())
    ops["Zm_P"] = csr_matrix(ops["Zm"] @ ops["P"])
    ops["Zp_P"] = csr_matrix(ops["Zp"] @ ops["P"])
    ops["P_Zm_dag"] = csr_matrix(ops["P"] @ ops["Zm_dag"])
    ops["P_Zp_dag"] = csr_matrix(ops["P"] @ ops["Zp_dag"])
    return ops


def _k(ops, names):
    out = ops[names[0]]
    for nme in names[1:]:
        out = kron(out, ops[nme], format="csr")
    return csr_matrix(out)


def dressed_site_operators(n):
    """Zn_dressed_site_operators(n, pure_theory=True) (Zn_operators.py:204-266)."""
    r = rishon_operators(n)
    ops = {}
    for op in ["n", "E", "E_square"]:
        ops[f"{op}_mx"] = _k(r, [op, "IDz", "IDz", "IDz"])
        ops[f"{op}_my"] = _k(r, ["IDz", op, "IDz", "IDz"])
        ops[f"{op}_px"] = _k(r, ["IDz", "IDz", op, "IDz"])
        ops[f"{op}_py"] = _k(r, ["IDz", "IDz", "IDz", op])
    ops["C_px,py"] = _k(r, ["IDz", "IDz", "Zm_P", "Zp_dag"])
    ops["C_py,mx"] = _k(r, ["P_Zp_dag", "P", "P", "Zm"])
    ops["C_mx,my"] = _k(r, ["Zm_P", "Zp_dag", "IDz", "IDz"])
    ops["C_my,px"] = _k(r, ["IDz", "Zm_P", "Zp_dag", "IDz"])
    E2 = 0
    for s in ["mx", "my", "px", "py"]:
        E2 = E2 + 0.5 * ops[f"E_square_{s}"]
    ops["E_square"] = csr_matrix(E2)
    return ops


def border_labels(cfg):
    """Zn_border_configs for the pure theory (Zn_operators.py:385-437)."""
    lab = []
    mx, my, px, py = [cfg[i] == 0 for i in range(4)]
    if mx: lab.append("mx")
    if my: lab.append("my")
    if px: lab.append("px")
    if py: lab.append("py")
    if mx and my: lab.append("mx_my")
    if mx and py: lab.append("mx_py")
    if my and px: lab.append("px_my")
    if px and py: lab.append("px_py")
    return lab


def gauge_invariant_states(n):
    """Zn_gauge_invariant_states(n, True, 2) (Zn_operators.py:269-360)."""
    borders = ["mx", "px", "my", "py", "mx_my", "mx_py", "px_my", "px_py"]
    rows = {"site": []}
    states = {"site": []}
    for b in borders:
        rows[f"site_{b}"] = []
        states[f"site_{b}"] = []
    D = n ** 4
    for row, cfg in enumerate(product(range(n), repeat=4)):
        if (sum(cfg) - 2 * (n - 1)) % n == 0:
            rows["site"].append(row)
            states["site"].append(cfg)
            for b in border_labels(cfg):
                rows[f"site_{b}"].append(row)
                states[f"site_{b}"].append(cfg)
    basis = {k: csr_matrix((np.ones(len(v)), (np.array(v, dtype=int), np.arange(len(v)))), shape=(D, len(v)))
             for k, v in rows.items()}
    return basis, {k: np.asarray(v) for k, v in states.items()}


def clock_operator(n, label):
    """E_clock on the full dressed site: sum over the half-links that belong to a
    lattice link (not the outer half-links of an OBC border label) of
    cos(2 pi E_h / n); summed over sites it is sum_links (E + E^dag)
    (each link = two half-links with E_p(i) = -E_m(j))."""
    r = rishon_operators(n)
    e = np.arange(n) - 0.5 * (n - 1)
    c = csr_matrix(diags(np.cos(2 * np.pi * e / n), 0, (n, n)))
    outer = set(label.split("_")[1:])
    out = 0
    for k, s in enumerate(["mx", "my", "px", "py"]):
        if s in outer:
            continue
        names = ["IDz"] * 4
        r2 = dict(r)
        r2["c"] = c
        names[k] = "c"
        out = out + _k(r2, names)
    return csr_matrix(out)
