"""Dressed-site local bases and operators, evaluated directly on the
gauge-invariant states.

A dressed site is a tuple of degrees of freedom (optional matter occupation,
then the half-links m_x, m_y, (m_z), p_x, p_y, (p_z)).  Every operator the
model needs is a signed product of single-dof matrices, one per dof (the
reference builds the same operators as Kronecker products and then projects
them, edlgt/operators/QED_operators.py:120-330 and
edlgt/models/quantum_model.py:400-502).  Here the matrix element between two
gauge-invariant states a, b is evaluated directly as
    coeff * prod_k F_k[a_k, b_k],
so no full-space Kronecker product is ever formed.

Local state order follows itertools.product over the dofs, E ascending from
-spin (QED_operators.py:449-485), and a label keeps the states satisfying the
Gauss law (and, for OBC border labels, a zero outer half-link).
"""
from itertools import product

import numpy as np


# ---------------------------------------------------------------- single-dof matrices
def rishon_matrices(spin):
    """U(1) rishon matrices in the E basis, E = -spin..spin (QED_operators.py:26-117,
    U='ladder', fermionic rishons)."""
    D = int(2 * spin + 1)
    P = np.diag([(-1.0) ** k for k in range(D)])
    U = np.zeros((D, D))
    for k in range(D - 1):
        U[k + 1, k] = 1.0  # raises E by one unit
    Zp = U @ P
    m = {
        "Iz": np.eye(D), "P": P, "U": U, "Zp": Zp, "Zm_dag": U, "Zp_dag": Zp.T, "Zm": U.T.copy(),
        "E": np.diag(np.arange(-spin, spin + 1, dtype=float)),
    }
    m["E2"] = m["E"] @ m["E"]
    m["Zm_P"] = m["Zm"] @ P
    m["Zp_P"] = Zp @ P
    m["P_Zm_dag"] = P @ U
    m["P_Zp_dag"] = P @ Zp.T
    return m


def matter_matrices():
    """Spinless fermion (bose_fermi_operators.py:14-70, has_spin=False)."""
    psi = np.array([[0.0, 1.0], [0.0, 0.0]])
    Pp = np.diag([1.0, -1.0])
    return {
        "ID": np.eye(2), "psi": psi, "psi_dag": psi.T.copy(), "P_psi": Pp,
        "N": np.diag([0.0, 1.0]), "N_zero": np.diag([1.0, 0.0]), "psi_dag_P": psi.T @ Pp,
    }


# ---------------------------------------------------------------- local bases
class DressedSite:
    """Gauge-invariant local states of one model, per site label.

    states[label]: int array (d_loc, n_dof) of dof indices;
    values[label]: the same states as physical values (n, E...)."""

    def __init__(self, dofs, labels):
        self.dofs = dofs          # list of (name, size)
        self.states = {}
        self.values = {}
        for lab, (idx, vals) in labels.items():
            self.states[lab] = np.asarray(idx, dtype=np.int64).reshape(-1, len(dofs))
            self.values[lab] = np.asarray(vals, dtype=np.int64).reshape(-1, len(dofs))

    def dim(self, label):
        return self.states[label].shape[0]

    def matrix(self, label, factors, coeff=1.0):
        """coeff * prod_k F_k[a_k, b_k] over the label's states."""
        st = self.states[label]
        M = np.full((st.shape[0], st.shape[0]), complex(coeff))
        for k, F in enumerate(factors):
            M = M * F[st[:, k][:, None], st[:, k][None, :]]
        return M


def _border_parts(dim):
    axes = "xyz"[:dim]
    if dim == 1:
        return [["mx"], ["px"]]
    if dim == 2:
        return [["mx"], ["px"], ["my"], ["py"], ["mx", "my"], ["mx", "py"], ["px", "my"], ["px", "py"]]
    parts = [[f"{s}{a}"] for a in axes for s in "mp"]
    for a, b in [("x", "y"), ("x", "z"), ("y", "z")]:
        for sa in "mp":
            for sb in "mp":
                parts.append([f"{sa}{a}", f"{sb}{b}"])
    for sx in "mp":
        for sy in "mp":
            for sz in "mp":
                parts.append([f"{sx}x", f"{sy}y", f"{sz}z"])
    return parts


def bg_label(label, q):
    """Label of a site whose static background charge is q (q = 0: the plain label)."""
    return label if q == 0 else f"{label}@bg{int(q)}"


def qed_site(spin, pure_theory, dim, backgrounds=(0,)):
    """Gauss-law states: sum_mu (E_p - E_m) = n - (0 even | 1 odd) + q with
    matter, = q without, for every static background charge q in backgrounds
    (labels of q != 0 carry the suffix "@bg<q>"); border labels additionally
    fix the named half-links to E=0 (QED_operators.py:410-538,
    lattice_geometry.py:345-487).  The reference makes the background a
    leading dof of the dressed site and keeps, per site, the columns of its
    own charge (quantum_model.py:440-466): the kept states are these, in the
    same product order."""
    axes = "xyz"[:dim]
    E_vals = list(range(-spin, spin + 1))
    dofs = ([] if pure_theory else [("n", 2)]) + [(f"E_m{a}", len(E_vals)) for a in axes] + \
           [(f"E_p{a}", len(E_vals)) for a in axes]
    cores = [("site", 0)] if pure_theory else [("even", 0), ("odd", 1)]
    off = 0 if pure_theory else 1
    link_pos = {f"m{a}": off + i for i, a in enumerate(axes)}
    link_pos.update({f"p{a}": off + dim + i for i, a in enumerate(axes)})
    labels = {}
    for q in sorted(set(int(b) for b in backgrounds)):
        for core, odd in cores:
            sets = {core: ([], [])}
            for parts in _border_parts(dim):
                sets[core + "_" + "_".join(parts)] = ([], [])
            ranges = [range(s) for _, s in dofs]
            for idx in product(*ranges):
                val = list(idx)
                for k in range(off, len(dofs)):
                    val[k] = E_vals[idx[k]]
                charge = (0 if pure_theory else val[0] - odd) + q
                div = sum(val[off + dim + m] - val[off + m] for m in range(dim))
                if div != charge:
                    continue
                for lab, (ii, vv) in sets.items():
                    parts = lab.split("_")[1:]
                    if all(val[link_pos[p]] == 0 for p in parts):
                        ii.append(idx)
                        vv.append(val)
            for lab, v in sets.items():
                if not v[0]:
                    continue  # no state of this label carries the charge q
                labels[bg_label(lab, q)] = v
    return DressedSite(dofs, labels)


def qed_operator_specs(pure_theory, dim):
    """Operator name -> list of (coeff, [factor name per dof]) terms
    (QED_operators.py:120-330)."""
    axes = "xyz"[:dim]
    nd = 2 * dim
    m0 = [] if pure_theory else ["ID"]

    def one(pos, name):
        f = ["Iz"] * nd
        f[pos] = name
        return m0 + f

    pos = {f"m{a}": i for i, a in enumerate(axes)}
    pos.update({f"p{a}": dim + i for i, a in enumerate(axes)})
    specs = {}
    for op in ["E", "E2", "P"]:
        for side in ["m", "p"]:
            for a in axes:
                specs[f"{op}_{side}{a}"] = [(1.0, one(pos[f"{side}{a}"], op))]
    specs["E2"] = [(0.5, one(pos[f"{s}{a}"], "E2")) for a in axes for s in "mp"]
    if dim >= 2:
        def corner(sign, assign):
            f = ["Iz"] * nd
            for k, v in assign.items():
                f[pos[k]] = v
            return [(sign, m0 + f)]

        if dim == 2:
            specs["C_px,py"] = corner(-1.0, {"px": "Zp_P", "py": "Zp_dag"})
            specs["C_py,mx"] = corner(1.0, {"mx": "P_Zm_dag", "my": "P", "px": "P", "py": "Zp"})
            specs["C_mx,my"] = corner(1.0, {"mx": "Zm_P", "my": "Zm_dag"})
            specs["C_my,px"] = corner(1.0, {"my": "Zm_P", "px": "Zp_dag"})
        else:
            specs["C_px,py"] = corner(-1.0, {"px": "Zp_P", "py": "Zp_dag"})
            specs["C_py,mx"] = corner(1.0, {"mx": "P_Zm_dag", "my": "P", "mz": "P", "px": "P", "py": "Zp"})
            specs["C_mx,my"] = corner(1.0, {"mx": "Zm_P", "my": "Zm_dag"})
            specs["C_my,px"] = corner(1.0, {"my": "Zm_P", "mz": "P", "px": "Zp_dag"})
            specs["C_px,pz"] = corner(-1.0, {"px": "Zp_P", "py": "P", "pz": "Zp_dag"})
            specs["C_pz,mx"] = corner(1.0, {"mx": "P_Zm_dag", "my": "P", "mz": "P", "px": "P", "py": "P", "pz": "Zp"})
            specs["C_mx,mz"] = corner(1.0, {"mx": "Zm_P", "my": "P", "mz": "Zm_dag"})
            specs["C_mz,px"] = corner(1.0, {"mz": "Zm_P", "px": "Zp_dag"})
            specs["C_py,pz"] = corner(-1.0, {"py": "Zp_P", "pz": "Zp_dag"})
            specs["C_pz,my"] = corner(1.0, {"my": "P_Zm_dag", "mz": "P", "px": "P", "py": "P", "pz": "Zp"})
            specs["C_my,mz"] = corner(1.0, {"my": "Zm_P", "mz": "Zm_dag"})
            specs["C_mz,py"] = corner(1.0, {"mz": "Zm_P", "px": "P", "py": "Zp_dag"})
    if not pure_theory:
        # 1D/2D hoppings carry the matter parity, 3D ones do not (QED_operators.py:172-175, 198-201, 244-262)
        mdag = "psi_dag_P" if dim < 3 else "psi_dag"
        for a in axes:
            f = ["Iz"] * nd
            f[pos[f"m{a}"]] = "Zm"
            for b in axes[: axes.index(a)]:
                f[pos[f"m{b}"]] = "P"
            specs[f"Q_m{a}_dag"] = [(1.0, [mdag] + f)]
            f = ["P"] * dim + ["Iz"] * dim
            f[pos[f"p{a}"]] = "Zp"
            for b in axes[: axes.index(a)]:
                f[pos[f"p{b}"]] = "P"
            specs[f"Q_p{a}_dag"] = [(1.0, [mdag] + f)]
        specs["N"] = [(1.0, ["N"] + ["Iz"] * nd)]
        specs["N_zero"] = [(1.0, ["N_zero"] + ["Iz"] * nd)]
    return specs


def qed_operators(site, spin, pure_theory, dim, label, with_bg=False):
    """All QED dressed-site operators on one label's gauge-invariant states
    (with_bg: also the background charge operator "bg")."""
    mats = rishon_matrices(spin)
    if not pure_theory:
        mats.update(matter_matrices())
    out = {}
    for name, terms in qed_operator_specs(pure_theory, dim).items():
        M = 0
        for coeff, fac in terms:
            M = M + site.matrix(label, [mats[f] for f in fac], coeff)
        out[name] = M
    if not pure_theory:
        for a in "xyz"[:dim]:
            for s in "mp":
                out[f"Q_{s}{a}"] = out[f"Q_{s}{a}_dag"].conj().T.copy()
    if with_bg:
        # the background charge operator (QED_operators.py:318-326): q on every state
        q = float(label.split("@bg")[1]) if "@bg" in label else 0.0
        out["bg"] = q * np.eye(site.dim(label))
    if dim == 3:
        out["EzC_px,py"] = 1j * (out["E_pz"] + out["E_mz"]) @ out["C_px,py"]
        out["EyC_px,pz"] = -1j * (out["E_py"] + out["E_my"]) @ out["C_px,pz"]
        out["ExC_py,pz"] = 1j * (out["E_px"] + out["E_mx"]) @ out["C_py,pz"]
    return out


# ---------------------------------------------------------------- Z_n (pure gauge, 2D)
def zn_rishon_matrices(n):
    """Z_n link/rishon matrices, pure theory (Zn_operators.py:130-201): U is the
    cyclic raising shift (U|k> = |k+1 mod n>), E = n_op - (n-1)/2, P = 1."""
    U = np.zeros((n, n))
    for k in range(n):
        U[(k + 1) % n, k] = 1.0
    E = np.diag(np.arange(n, dtype=float) - 0.5 * (n - 1))
    return {"IDz": np.eye(n), "U": U, "Ut": U.T.copy(), "n": np.diag(np.arange(n, dtype=float)),
            "E": E, "E_square": E @ E}


def zn_site(n, dim=2):
    """Gauss law sum(rishon indices) = dim*(n-1) (mod n) over (m_x, m_y, p_x, p_y)
    (Zn_operators.py:269-360); OBC border labels require index 0 on the named
    half-links (Zn_border_configs, Zn_operators.py:385-437)."""
    if dim != 2:
        raise NotImplementedError("Z_n dressed sites are defined for 2D lattices")
    dofs = [(f"n_{s}", n) for s in ("mx", "my", "px", "py")]
    pos = {"mx": 0, "my": 1, "px": 2, "py": 3}
    sets = {"site": ([], [])}
    for parts in _border_parts(2):
        sets["site_" + "_".join(parts)] = ([], [])
    for idx in product(*[range(n)] * 4):
        if (sum(idx) - dim * (n - 1)) % n != 0:
            continue
        for lab, (ii, vv) in sets.items():
            parts = lab.split("_")[1:]
            if all(idx[pos[p]] == 0 for p in parts):
                ii.append(idx)
                vv.append(idx)
    return DressedSite(dofs, sets)


def zn_operators(site, n, label):
    """Dressed-site Z_n operators on one label (Zn_operators.py:204-266), plus the
    clock electric operator E_clock = sum over the label's linked half-links of
    cos(2 pi E_h / n) (so that summing it over sites gives sum_links (E + E^dag))."""
    m = zn_rishon_matrices(n)
    order = ["mx", "my", "px", "py"]

    def op(assign, coeff=1.0):
        return site.matrix(label, [m[assign.get(s, "IDz")] for s in order], coeff)

    out = {}
    for name in ("n", "E", "E_square"):
        for s in order:
            out[f"{name}_{s}"] = op({s: name})
    out["E_square"] = sum(0.5 * out[f"E_square_{s}"] for s in order)
    out["C_px,py"] = op({"px": "U", "py": "Ut"})
    out["C_py,mx"] = op({"mx": "Ut", "py": "U"})
    out["C_mx,my"] = op({"mx": "U", "my": "Ut"})
    out["C_my,px"] = op({"my": "U", "px": "Ut"})
    outer = set(label.split("_")[1:])
    clock = np.zeros_like(out["E_mx"])
    st = site.states[label]
    for k, s in enumerate(order):
        if s in outer:
            continue
        e = st[:, k] - 0.5 * (n - 1)
        clock = clock + np.diag(np.cos(2 * np.pi * e / n))
    out["E_clock"] = clock
    return out
