"""U(1) (QED) local operators and gauge-invariant bases (oracle; test infrastructure).

Restates edlgt/operators/QED_operators.py:26-538 and
edlgt/operators/bose_fermi_operators.py:14-70 with scipy Kronecker products,
exactly in the reference's factor order (static background charge first when
present, then matter, then the half-links m_x, m_y, ..., p_x, p_y, ...).
Only the ``U="ladder"``, fermionic branch used by the hot-path
configurations is restated.
"""
from itertools import product

import numpy as np
from scipy.sparse import csr_matrix, diags, identity, kron

from .lattice import border_configs, border_labels


def _kron_list(ops, names):
    """qmb_operator: kron of named single-factor matrices (qmb_operations.py:28-73)."""
    out = ops[names[0]]
    for n in names[1:]:
        out = kron(out, ops[n], format="csr")
    return csr_matrix(out)


def rishon_operators(spin):
    """QED_rishon_operators, ladder + fermionic branch (QED_operators.py:26-117)."""
    size = int(2 * spin + 1)
    shape = (size, size)
    ops = {}
    ops["P"] = diags([(-1.0) ** i for i in range(size)], 0, shape, dtype=float)
    U = diags(np.ones(size - 1), -1, shape, dtype=float)  # raises E by one unit
    ops["U"] = U
    ops["Zp"] = U @ ops["P"]
    ops["Zm_dag"] = U
    ops["Zp_dag"] = ops["Zp"].transpose()
    ops["Zm"] = ops["Zm_dag"].transpose()
    ops["Zm_P"] = ops["Zm"] @ ops["P"]
    ops["Zp_P"] = ops["Zp"] @ ops["P"]
    ops["P_Zm_dag"] = ops["P"] @ ops["Zm_dag"]
    ops["P_Zp_dag"] = ops["P"] @ ops["Zp_dag"]
    ops["Iz"] = identity(size, dtype=float)
    ops["E"] = diags(np.arange(-spin, spin + 1, 1, dtype=float), 0, shape)
    ops["E2"] = ops["E"] @ ops["E"]
    return {k: csr_matrix(v) for k, v in ops.items()}


def fermi_operators():
    """Spinless fermion matter (bose_fermi_operators.py:14-70, has_spin=False)."""
    ops = {}
    ops["psi"] = csr_matrix(diags(np.array([1.0]), 1, (2, 2)))
    ops["psi_dag"] = csr_matrix(ops["psi"].transpose())
    ops["P_psi"] = csr_matrix(diags(np.array([1.0, -1.0]), 0, (2, 2)))
    ops["N"] = csr_matrix(ops["psi_dag"] @ ops["psi"])
    ops["ID_psi"] = csr_matrix(identity(2))
    ops["ID"] = csr_matrix(identity(2))
    ops["psi_dag_P"] = csr_matrix(ops["psi_dag"] @ ops["P_psi"])
    ops["N_zero"] = csr_matrix(ops["ID"] - ops["N"])
    return ops


def dressed_site_operators(spin, pure_theory, lattice_dim, background=0):
    """QED_dressed_site_operators for 1D/2D/3D (QED_operators.py:120-330);
    background > 0 prepends the static-charge dof (values -background..
    background) to every operator as an identity factor and adds its diagonal
    operator "bg" (QED_operators.py:318-326)."""
    r = rishon_operators(spin)
    in_ops = dict(r)
    if not pure_theory:
        in_ops.update(fermi_operators())
    ops = {}
    if lattice_dim == 1:
        for op in ["E", "E2", "P"]:
            ops[f"{op}_mx"] = _kron_list(in_ops, [op, "Iz"])
            ops[f"{op}_px"] = _kron_list(in_ops, ["Iz", op])
        if not pure_theory:
            for k in list(ops):
                ops[k] = csr_matrix(kron(in_ops["ID_psi"], ops[k]))
            ops["Q_mx_dag"] = _kron_list(in_ops, ["psi_dag_P", "Zm", "Iz"])
            ops["Q_px_dag"] = _kron_list(in_ops, ["psi_dag_P", "P", "Zp"])
            for k in [k for k in ops if k.endswith("_dag")]:
                ops[k[:-4]] = csr_matrix(ops[k].conj().transpose())
            ops["N"] = _kron_list(in_ops, ["N", "Iz", "Iz"])
            ops["N_zero"] = _kron_list(in_ops, ["N_zero", "Iz", "Iz"])
    elif lattice_dim == 2:
        I4 = ["Iz"] * 4
        for op in ["E", "E2", "P"]:
            for pos, name in enumerate(["mx", "my", "px", "py"]):
                lst = list(I4)
                lst[pos] = op
                ops[f"{op}_{name}"] = _kron_list(in_ops, lst)
        ops["C_px,py"] = -_kron_list(in_ops, ["Iz", "Iz", "Zp_P", "Zp_dag"])
        ops["C_py,mx"] = _kron_list(in_ops, ["P_Zm_dag", "P", "P", "Zp"])
        ops["C_mx,my"] = _kron_list(in_ops, ["Zm_P", "Zm_dag", "Iz", "Iz"])
        ops["C_my,px"] = _kron_list(in_ops, ["Iz", "Zm_P", "Zp_dag", "Iz"])
        if not pure_theory:
            for k in list(ops):
                ops[k] = csr_matrix(kron(in_ops["ID_psi"], ops[k]))
            ops["Q_mx_dag"] = _kron_list(in_ops, ["psi_dag_P", "Zm", "Iz", "Iz", "Iz"])
            ops["Q_my_dag"] = _kron_list(in_ops, ["psi_dag_P", "P", "Zm", "Iz", "Iz"])
            ops["Q_px_dag"] = _kron_list(in_ops, ["psi_dag_P", "P", "P", "Zp", "Iz"])
            ops["Q_py_dag"] = _kron_list(in_ops, ["psi_dag_P", "P", "P", "P", "Zp"])
            for k in [k for k in ops if k.endswith("_dag")]:
                ops[k[:-4]] = csr_matrix(ops[k].conj().transpose())
            ops["N"] = _kron_list(in_ops, ["N", "Iz", "Iz", "Iz", "Iz"])
            ops["N_zero"] = _kron_list(in_ops, ["N_zero", "Iz", "Iz", "Iz", "Iz"])
    elif lattice_dim == 3:
        names6 = ["mx", "my", "mz", "px", "py", "pz"]
        for op in ["E", "E2", "P"]:
            for pos, name in enumerate(names6):
                lst = ["Iz"] * 6
                lst[pos] = op
                ops[f"{op}_{name}"] = _kron_list(in_ops, lst)
        corners = {
            "C_px,py": (-1, ["Iz", "Iz", "Iz", "Zp_P", "Zp_dag", "Iz"]),
            "C_py,mx": (1, ["P_Zm_dag", "P", "P", "P", "Zp", "Iz"]),
            "C_mx,my": (1, ["Zm_P", "Zm_dag", "Iz", "Iz", "Iz", "Iz"]),
            "C_my,px": (1, ["Iz", "Zm_P", "P", "Zp_dag", "Iz", "Iz"]),
            "C_px,pz": (-1, ["Iz", "Iz", "Iz", "Zp_P", "P", "Zp_dag"]),
            "C_pz,mx": (1, ["P_Zm_dag", "P", "P", "P", "P", "Zp"]),
            "C_mx,mz": (1, ["Zm_P", "P", "Zm_dag", "Iz", "Iz", "Iz"]),
            "C_mz,px": (1, ["Iz", "Iz", "Zm_P", "Zp_dag", "Iz", "Iz"]),
            "C_py,pz": (-1, ["Iz", "Iz", "Iz", "Iz", "Zp_P", "Zp_dag"]),
            "C_pz,my": (1, ["Iz", "P_Zm_dag", "P", "P", "P", "Zp"]),
            "C_my,mz": (1, ["Iz", "Zm_P", "Zm_dag", "Iz", "Iz", "Iz"]),
            "C_mz,py": (1, ["Iz", "Iz", "Zm_P", "P", "Zp_dag", "Iz"]),
        }
        for name, (sg, lst) in corners.items():
            ops[name] = csr_matrix(sg * _kron_list(in_ops, lst))
        ops["EzC_px,py"] = csr_matrix(1j * (ops["E_pz"] + ops["E_mz"]) @ ops["C_px,py"])
        ops["EyC_px,pz"] = csr_matrix(-1j * (ops["E_py"] + ops["E_my"]) @ ops["C_px,pz"])
        ops["ExC_py,pz"] = csr_matrix(1j * (ops["E_px"] + ops["E_mx"]) @ ops["C_py,pz"])
        if not pure_theory:
            for k in list(ops):
                ops[k] = csr_matrix(kron(in_ops["ID_psi"], ops[k]))
            hops = {"Q_mx_dag": ["psi_dag", "Zm", "Iz", "Iz", "Iz", "Iz", "Iz"],
                    "Q_my_dag": ["psi_dag", "P", "Zm", "Iz", "Iz", "Iz", "Iz"],
                    "Q_mz_dag": ["psi_dag", "P", "P", "Zm", "Iz", "Iz", "Iz"],
                    "Q_px_dag": ["psi_dag", "P", "P", "P", "Zp", "Iz", "Iz"],
                    "Q_py_dag": ["psi_dag", "P", "P", "P", "P", "Zp", "Iz"],
                    "Q_pz_dag": ["psi_dag", "P", "P", "P", "P", "P", "Zp"]}
            for k, lst in hops.items():
                ops[k] = _kron_list(in_ops, lst)
            for k in [k for k in ops if k.endswith("_dag")]:
                ops[k[:-4]] = csr_matrix(ops[k].conj().transpose())
            ops["N"] = _kron_list(in_ops, ["N"] + ["Iz"] * 6)
            ops["N_zero"] = _kron_list(in_ops, ["N_zero"] + ["Iz"] * 6)
    else:
        raise NotImplementedError("oracle covers 1D, 2D and 3D QED")
    E2 = 0
    for d in "xyz"[:lattice_dim]:
        for s in "mp":
            E2 = E2 + 0.5 * ops[f"E2_{s}{d}"]
    ops["E2"] = csr_matrix(E2)
    if background > 0:
        bg_vals = np.arange(-background, background + 1, dtype=float)
        I_bg = identity(bg_vals.size, dtype=float)
        for k in list(ops):
            ops[k] = csr_matrix(kron(I_bg, ops[k]))
        D = ops["E2"].shape[0] // bg_vals.size
        ops["bg"] = csr_matrix(kron(diags(bg_vals, 0), identity(D, dtype=float)))
    return ops


def gauge_invariant_states(spin, pure_theory, lattice_dim, background=0):
    """QED_gauge_invariant_states (QED_operators.py:410-538).

    Returns (basis, states): basis[label] is the (D_full, d_loc) 0/1 isometry,
    states[label] the (d_loc, n_dof) local configurations, in itertools.product
    order over (bg, n, E_m..., E_p...) with E ascending from -spin; the static
    charge bg (values -background..background, present when background > 0)
    enters the Gauss law as div E = matter charge + bg and is ignored by the
    border classification."""
    rish = list(range(-spin, spin + 1))
    dof = [rish] * (2 * lattice_dim)
    if pure_theory:
        cores, parity = ["site"], [1]
    else:
        cores, parity = ["even", "odd"], [1, -1]
        dof = [[0, 1]] + dof
    if background > 0:
        dof = [list(range(-background, background + 1))] + dof
    borders = border_labels(lattice_dim)
    rows, states = {}, {}
    D_full = 0
    for ii, core in enumerate(cores):
        keys = [core] + [f"{core}_{b}" for b in borders]
        for k in keys:
            rows[k], states[k] = [], []
        for row, full in enumerate(product(*dof)):
            D_full = row + 1
            q_bg, cfg = (full[0], full[1:]) if background > 0 else (0, full)
            if pure_theory:
                charge, links = q_bg, cfg
            else:
                charge = cfg[0] - (0 if parity[ii] > 0 else 1) + q_bg
                links = cfg[1:]
            div = sum(links[mu + lattice_dim] - links[mu] for mu in range(lattice_dim))
            if div == charge:
                rows[core].append(row)
                states[core].append(full)
                for b in border_configs(list(links), lattice_dim, 0):
                    rows[f"{core}_{b}"].append(row)
                    states[f"{core}_{b}"].append(full)
    basis = {}
    for k in rows:
        n = len(rows[k])
        basis[k] = csr_matrix(
            (np.ones(n), (np.array(rows[k], dtype=int), np.arange(n))), shape=(D_full, n)
        )
        states[k] = np.asarray(states[k])
    return basis, states
