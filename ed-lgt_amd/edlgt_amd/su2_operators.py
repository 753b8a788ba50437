"""SU(2) dressed sites: rishons of every irrep j <= s_max per half-link, with
optional two-colour staggered matter.

A dressed site is matter (4 states: vacuum, r = m +1/2, g = m -1/2, pair)
times one rishon per half-link (the irreps j = 0, 1/2, ..., s_max, each with
m = j, ..., -j: 3 states for s_max = 1/2), half-links ordered m_x, (m_y,
m_z,) p_x, (p_y, p_z) (edlgt/operators/SU2_operators.py:35-300, 302-592,
SU2_rishons.py:50-234, bose_fermi_operators.py:14-70).  The gauge-invariant
local states are the SU(2) singlets of the site's spins, built by coupling
the spins one after the other (matter first) and grouped by their
intermediate total spins (SU2_singlets.py:30-249); their order follows
SU2_gauge_invariant_states (SU2_operators.py:625-714): spin configurations in
product order, then singlets sorted by intermediate spins.  Operators are
sparse Kronecker products on the full site space projected onto each
label's singlet basis, B^T O B (quantum_model.py:400-502).

Two operator sets: the hard-coded spin-1/2 one (SU2_dressed_site_operators,
rishons Z_r, Z_g that raise and lower j) and the generalised one of any
s_max (SU2_gen_dressed_site_operators: rishons that lower j by 1/2, in two
species A and B, hoppings Q1/Q2 and sixteen corner combinations per plaquette).

Clebsch-Gordan coefficients come from the Racah formula in exact rational
arithmetic (a coefficient is sign * sqrt(rational); a singlet amplitude, a
product of them, is rounded once).
"""
from fractions import Fraction
from itertools import product
from math import factorial, sqrt

import numpy as np
import scipy.sparse as sp

HALF = Fraction(1, 2)


# ---------------------------------------------------------------- Clebsch-Gordan
def _fact(x):
    if x.denominator != 1 or x < 0:
        raise ValueError("non-integer factorial argument")
    return factorial(int(x))


def cg_exact(j1, m1, j2, m2, J, M):
    """<j1 m1 j2 m2 | J M> as (sign_rational, radicand) with value
    sign_rational * sqrt(radicand); arguments are Fractions."""
    if m1 + m2 != M or abs(m1) > j1 or abs(m2) > j2 or abs(M) > J:
        return Fraction(0), Fraction(0)
    if J < abs(j1 - j2) or J > j1 + j2:
        return Fraction(0), Fraction(0)
    for a in (j1 + m1, j2 + m2, J + M, j1 + j2 + J):
        if a.denominator != 1:
            return Fraction(0), Fraction(0)
    rad = Fraction((2 * J + 1).numerator, (2 * J + 1).denominator)
    rad *= Fraction(_fact(J + j1 - j2) * _fact(J - j1 + j2) * _fact(j1 + j2 - J), _fact(j1 + j2 + J + 1))
    rad *= _fact(J + M) * _fact(J - M) * _fact(j1 - m1) * _fact(j1 + m1) * _fact(j2 - m2) * _fact(j2 + m2)
    s = Fraction(0)
    k = 0
    while True:
        args = [j1 + j2 - J - k, j1 - m1 - k, j2 + m2 - k, J - j2 + m1 + k, J - j1 - m2 + k]
        if any(a < 0 for a in args[:3]):
            break
        if all(a >= 0 for a in args):
            den = factorial(k)
            for a in args:
                den *= _fact(Fraction(a))
            s += Fraction((-1) ** k, den)
        k += 1
    return s, rad


def m_values(j):
    """m = j, j-1, ..., -j (descending, spin_operators.py:25-28)."""
    return [j - k for k in range(int(2 * j) + 1)]


def _couple(j1, j2, singlet, m1_fixed=None):
    """Non-zero couplings of j1 (all m1, or m1 = m1_fixed) with j2, in the
    reference's enumeration order: m1, then m2 (descending), then J
    ascending (only J = 0 for a singlet).  Entries (m1, m2, cg, J, M)."""
    m1s = [m1_fixed] if m1_fixed is not None else m_values(j1)
    Js = [Fraction(0)] if singlet else [abs(j1 - j2) + k for k in range(int(j1 + j2 - abs(j1 - j2)) + 1)]
    out = []
    for m1 in m1s:
        for m2 in m_values(j2):
            for J in Js:
                M = m1 + m2
                if abs(M) > J:
                    continue
                s, r = cg_exact(j1, m1, j2, m2, J, M)
                if s != 0:
                    out.append((m1, m2, (s, r), J, M))
    return out


def su2_singlets(spins):
    """Singlets of the spins (Fractions) coupled left to right.  Returns a list
    of (intermediate_J_tuple, [(m_tuple, amplitude_float), ...]) in the
    reference's order (stable sort by intermediate spins)."""
    n = len(spins)
    if n < 2:
        raise ValueError("at least two spins are needed")
    chains = []  # (m list, (sign, rad) product, J list, M)
    for m1, m2, c, J, M in _couple(spins[0], spins[1], n == 2):
        chains.append(([m1, m2], c, [J], M))
    for k in range(2, n):
        nxt = []
        for ms, (s, r), Js, M in chains:
            for _, m2, (s2, r2), J, M2 in _couple(Js[-1], spins[k], k == n - 1, m1_fixed=M):
                nxt.append((ms + [m2], (s * s2, r * r2), Js + [J], M2))
        chains = nxt
    if not chains:
        return []
    chains = sorted(chains, key=lambda c: tuple(c[2]))  # stable, like the reference
    groups = []
    for ms, (s, r), Js, _ in chains:
        amp = float(s) * sqrt(float(r))
        if groups and groups[-1][0] == tuple(Js):
            groups[-1][1].append((tuple(ms), amp))
        else:
            groups.append((tuple(Js), [(tuple(ms), amp)]))
    return groups


# ---------------------------------------------------------------- single-dof matrices
def rishon_matrices():
    """Spin-1/2 rishon algebra (SU2_rishons.py:50-80) and generators
    (spin_operators.py:91-127): basis (j=0), (1/2, +1/2), (1/2, -1/2)."""
    cf = 2.0 ** -0.25
    P = np.diag([1.0, -1.0, -1.0])
    Zr = np.zeros((3, 3))
    Zr[0, 1] = cf
    Zr[2, 0] = cf
    Zg = np.zeros((3, 3))
    Zg[0, 2] = cf
    Zg[1, 0] = -cf
    Tz = np.diag([0.0, 0.5, -0.5])
    Tp = np.zeros((3, 3))
    Tp[1, 2] = 1.0
    Tm = Tp.T.copy()
    m = {"Iz": np.eye(3), "P": P, "Zr": Zr, "Zg": Zg, "Tz": Tz, "Tp": Tp, "Tm": Tm,
         "Tx": 0.5 * (Tp + Tm), "Ty": -0.5j * (Tp - Tm), "T2": np.diag([0.0, 0.75, 0.75])}
    for c in "rg":
        Z = m[f"Z{c}"]
        m[f"Z{c}_dag"] = Z.T.copy()
        m[f"Z{c}_P"] = Z @ P
        m[f"P_Z{c}_dag"] = P @ Z.T
        m[f"Z{c}_dag_P"] = Z.T @ P
    return m


def matter_matrices():
    """Two-component fermion on one site (bose_fermi_operators.py:43-67):
    index = 2 * n_down + n_up, r = up (m = +1/2), g = down (m = -1/2)."""
    psi = np.array([[0.0, 1.0], [0.0, 0.0]])
    Pp = np.diag([1.0, -1.0])
    I2 = np.eye(2)
    N = np.diag([0.0, 1.0])
    m = {"ID_psi": np.eye(4), "P_psi": np.kron(Pp, Pp)}
    m["psi_g"] = np.kron(psi, I2)          # down
    m["psi_r"] = np.kron(Pp, psi)          # up, with the parity string of the down mode
    m["N_g"] = np.kron(N, I2)
    m["N_r"] = np.kron(I2, N)
    m["N_pair"] = m["N_r"] @ m["N_g"]
    m["N_tot"] = m["N_r"] + m["N_g"]
    m["N_single"] = m["N_tot"] - 2 * m["N_pair"]
    m["N_zero"] = np.eye(4) - m["N_pair"] - m["N_single"]
    for c in "rg":
        m[f"psi_{c}_dag"] = m[f"psi_{c}"].T.copy()
        m[f"psi_{c}_dag_P"] = m[f"psi_{c}_dag"] @ m["P_psi"]
    return m


def link_irreps(s_max):
    """j = 0, 1/2, ..., s_max."""
    return [Fraction(k, 2) for k in range(int(2 * s_max) + 1)]


def rishon_states(s_max):
    """(j, m) of every rishon index: irreps in increasing j, m descending."""
    return [(j, m) for j in link_irreps(s_max) for m in m_values(j)]


def generators(s_max):
    """Block-diagonal SU(2) generators on the rishon space (spin_operators.py:
    75-127, get_spin_operators :31-60): Tz = m, Tp |j m> = sqrt(j(j+1) -
    m(m+1)) |j m+1>, T2 = j(j+1), parity P = (-1)^(2j)."""
    st = rishon_states(s_max)
    n = len(st)
    Tz = np.diag([float(m) for _, m in st])
    T2 = np.diag([float(j * (j + 1)) for j, _ in st])
    P = np.diag([(-1.0) ** int(2 * j) for j, _ in st])
    Tp = np.zeros((n, n))
    for a, (ja, ma) in enumerate(st):
        for b, (jb, mb) in enumerate(st):
            if ja == jb and ma == mb + 1:
                Tp[a, b] = sqrt(float(jb * (jb + 1) - mb * (mb + 1)))
    Tm = Tp.T.copy()
    return {"Iz": np.eye(n), "P": P, "Tz": Tz, "Tp": Tp, "Tm": Tm, "Tx": 0.5 * (Tp + Tm),
            "Ty": -0.5j * (Tp - Tm), "T2": T2}


def _chi(j, color, m):
    """Coefficient of the generalised rishons (SU2_rishons.py:208-234)."""
    num = float(j + m + 1) if color == "r" else float(j - m + 1)
    return sqrt(num / sqrt(float((2 * j + 1) * (2 * j + 2))))


def rishon_matrices_gen(s_max):
    """Generalised rishons (SU2_Rishon_gen, SU2_rishons.py:106-196): Z_r maps
    |j + 1/2, m + 1/2> to |j, m> and Z_g maps |j + 1/2, m - 1/2> to |j, m> with
    coefficient chi(j, color, m); species B = Z, species A: ZA_r = ZB_g^dag,
    ZA_g = -ZB_r^dag; products with the parity P for the corner operators."""
    st = rishon_states(s_max)
    idx = {jm: k for k, jm in enumerate(st)}
    n = len(st)
    m = generators(s_max)
    Z = {"r": np.zeros((n, n)), "g": np.zeros((n, n))}
    for (j, mm), a in idx.items():
        if j == s_max:
            continue
        Z["r"][a, idx[(j + HALF, mm + HALF)]] = _chi(j, "r", mm)
        Z["g"][a, idx[(j + HALF, mm - HALF)]] = _chi(j, "g", mm)
    P = m["P"]
    for c in "rg":
        m[f"Z{c}"] = Z[c]
        m[f"Z{c}_dag"] = Z[c].T.copy()
        m[f"ZB_{c}"] = Z[c]
        m[f"ZB_{c}_dag"] = Z[c].T.copy()
    m["ZA_r"] = m["ZB_g_dag"].copy()
    m["ZA_g"] = -m["ZB_r_dag"]
    for sp_ in "AB":
        for c in "rg":
            z = m[f"Z{sp_}_{c}"]
            m[f"Z{sp_}_{c}_dag"] = z.T.copy()
            m[f"Z{sp_}_{c}_P"] = z @ P
            m[f"P_Z{sp_}_{c}_dag"] = P @ z.T
    return m


def _kron(mats):
    out = sp.identity(1, format="csr")
    for a in mats:
        out = sp.kron(out, sp.csr_matrix(a), format="csr")
    return out


# ---------------------------------------------------------------- dressed site
class SU2Site:
    """Singlet bases per site label ('site' and OBC border labels) as
    isometries basis[label] (full site dim x d_loc), for rishon irreps up to
    s_max; generic=True selects the generalised operator set."""

    def __init__(self, pure_theory, dim, s_max=HALF, generic=False, background=0):
        self.pure = bool(pure_theory)
        self.dim_lat = int(dim)
        self.s_max = Fraction(s_max).limit_denominator(2)
        self.generic = bool(generic)
        # static background charge: a leading spin of irrep 0..background per
        # site, coupled first (SU2_operators.py:625-714, SU2_singlets.py:329-387);
        # labels get the suffix "@bg<2j>" of the site's charge
        self.bg = Fraction(background).limit_denominator(2)
        self.bg_irreps = link_irreps(self.bg) if self.bg > 0 else []
        self.bgdim = len(rishon_states(self.bg)) if self.bg > 0 else 1
        if not self.generic and self.s_max != HALF:
            raise ValueError("the hard-coded SU(2) operator set is the spin-1/2 one")
        axes = "xyz"[:dim]
        self.links = [f"m{a}" for a in axes] + [f"p{a}" for a in axes]
        n_link = len(self.links)
        self.rdim = len(rishon_states(self.s_max))
        self._roff = {}
        k = 0
        for j in link_irreps(self.s_max):
            self._roff[j] = k
            k += int(2 * j) + 1
        self.full_dim = self.bgdim * (1 if self.pure else 4) * self.rdim ** n_link
        from .operators import _border_parts
        base_labels = ["site"] + ["site_" + "_".join(p) for p in _border_parts(dim)]
        labels = base_labels if not self.bg_irreps else [f"{b}@bg{int(2 * j)}" for b in base_labels
                                                         for j in self.bg_irreps]
        cols = {lab: [] for lab in labels}
        vals = {lab: [] for lab in labels}
        link_choices = [link_irreps(self.s_max)] * n_link
        matter_choices = [] if self.pure else [["V", "S", "P"]]
        bg_choices = [self.bg_irreps] if self.bg_irreps else []
        for combo in product(*(bg_choices + matter_choices + link_choices)):
            combo = list(combo)
            jbg = combo.pop(0) if self.bg_irreps else None
            if self.pure:
                mkind, links = None, combo
                spins = list(links)
            else:
                mkind, links = combo[0], combo[1:]
                spins = [HALF if mkind == "S" else Fraction(0)] + list(links)
            if jbg is not None:
                spins = [jbg] + spins
            j2 = [int(2 * s) for s in spins]
            tot = sum(j2)
            if tot % 2 or max(j2) > tot - max(j2):
                continue
            for Js, terms in su2_singlets(spins):
                vec = np.zeros(self.full_dim)
                for ms, amp in terms:
                    vec[self._index(mkind, spins, ms, jbg)] = amp
                zero_links = {self.links[k] for k, s in enumerate(links) if s == 0}
                jv = ([int(2 * jbg)] if jbg is not None else []) + \
                     ([{"V": 0, "S": 1, "P": 2}[mkind]] if not self.pure else []) + [int(2 * s) for s in links]
                for b in base_labels:
                    parts = b.split("_")[1:]
                    if all(p in zero_links for p in parts):
                        lab = b if jbg is None else f"{b}@bg{int(2 * jbg)}"
                        cols[lab].append(vec)
                        vals[lab].append(jv)
        self.basis = {lab: np.column_stack(cols[lab]) for lab in labels if cols[lab]}
        self.values = {lab: np.asarray(vals[lab], dtype=np.int64) for lab in labels if cols[lab]}

    def _index(self, mkind, spins, ms, jbg=None):
        """Full-site index: background (if any, most significant), matter
        (vacuum, r, g, pair), then one rishon index per half-link."""
        idx = 0
        if jbg is not None:
            # background irreps 0..bg, m descending within each
            off = sum(int(2 * j) + 1 for j in self.bg_irreps if j < jbg)
            idx = off + int(jbg - ms[0])
            spins, ms = spins[1:], ms[1:]
        k0 = 0
        if not self.pure:
            if mkind == "V":
                mi = 0
            elif mkind == "P":
                mi = 3
            else:
                mi = 1 if ms[0] > 0 else 2
            idx = idx * 4 + mi
            k0 = 1
        for s_, m in zip(spins[k0:], ms[k0:]):
            idx = idx * self.rdim + self._roff[s_] + int(s_ - m)
        return idx

    def dim(self, label):
        return self.basis[label].shape[1]

    def _op_factory(self, R, Mt):
        links = self.links
        bg = generators(self.bg) if self.bg > 0 else None
        if bg is not None:
            # the background slot carries its generators and, for the noise
            # term, the spin-1/2 rishons (SU2_operators.py:272-298)
            bg = dict(bg)
            if self.bg == HALF:
                r12 = rishon_matrices()
                for c in "rg":
                    bg[f"Z{c}"] = r12[f"Z{c}"]
                    bg[f"Z{c}_dag"] = r12[f"Z{c}_dag"]

        def op(assign, matter="ID_psi", background="Iz"):
            mats = [] if bg is None else [bg[background]]
            mats += [] if self.pure else [Mt[matter]]
            mats += [R[assign.get(l, "Iz")] for l in links]
            return _kron(mats)
        return op

    def _add_background(self, ops, op):
        """The site's background Casimir and, in 1D with matter and charges of
        spin 1/2, the Gauss-law-violating hoppings V (SU2_operators.py:272-298)."""
        if self.bg == 0:
            return
        ops["bg"] = op({}, background="T2")
        if self.dim_lat == 1 and not self.pure and self.bg == HALF:
            ops["Vmx_dag"] = sum(op({"mx": f"Z{c}"}, matter="P_psi", background=f"Z{c}_dag") for c in "rg")
            ops["Vpx_dag"] = sum(op({"mx": "P", "px": f"Z{c}"}, matter="P_psi", background=f"Z{c}_dag")
                                 for c in "rg")
            ops["Vmx"] = ops["Vmx_dag"].conj().T.tocsr()
            ops["Vpx"] = ops["Vpx_dag"].conj().T.tocsr()

    def _corner_lists(self):
        """Corner operator factor lists per (corner, colour) as (half-link ->
        factor name) maps, for the planes of the lattice (SU2_operators.py:
        104-131, 200-257, 383-399, 431-492); the names use a species prefix
        ('' for the hard-coded set, 'A'/'B' pairs for the generalised one)."""
        if self.dim_lat == 2:
            return {"px,py": {"px": "{a}_P", "py": "{b}_dag"},
                    "py,mx": {"mx": "P_{b}_dag", "my": "P", "px": "P", "py": "{a}"},
                    "mx,my": {"mx": "{a}_P", "my": "{b}_dag"},
                    "my,px": {"my": "{a}_P", "px": "{b}_dag"}}
        if self.dim_lat == 3:
            return {"px,py": {"px": "{a}_P", "py": "{b}_dag"},
                    "py,mx": {"mx": "P_{b}_dag", "my": "P", "mz": "P", "px": "P", "py": "{a}"},
                    "mx,my": {"mx": "{a}_P", "my": "{b}_dag"},
                    "my,px": {"my": "{a}_P", "mz": "P", "px": "{b}_dag"},
                    "px,pz": {"px": "{a}_P", "py": "P", "pz": "{b}_dag"},
                    "pz,mx": {"mx": "P_{b}_dag", "my": "P", "mz": "P", "px": "P", "py": "P", "pz": "{a}"},
                    "mx,mz": {"mx": "{a}_P", "my": "P", "mz": "{b}_dag"},
                    "mz,px": {"mz": "{a}_P", "px": "{b}_dag"},
                    "py,pz": {"py": "{a}_P", "pz": "{b}_dag"},
                    "pz,my": {"my": "P_{b}_dag", "mz": "P", "px": "P", "py": "P", "pz": "{a}"},
                    "my,mz": {"my": "{a}_P", "mz": "{b}_dag"},
                    "mz,py": {"mz": "{a}_P", "px": "P", "py": "{b}_dag"}}
        return {}

    def full_operators(self):
        """Sparse Kronecker-product operators on the full site space."""
        return self._generic_operators() if self.generic else self._standard_operators()

    def _standard_operators(self):
        """SU2_dressed_site_operators (SU2_operators.py:35-300), spin 1/2, 1D-3D."""
        R = rishon_matrices()
        Mt = None if self.pure else matter_matrices()
        links = self.links
        op = self._op_factory(R, Mt)
        ops = {}
        for name in ("T2", "P", "Tx", "Ty", "Tz"):
            for l in links:
                ops[f"{name}_{l}"] = op({l: name})
        if self.dim_lat == 1:
            ops["U"] = sum(op({"mx": f"Z{c}_dag_P", "px": f"Z{c}"}) for c in "rg") / np.sqrt(2)
        for corner, fac in self._corner_lists().items():
            ops[f"C_{corner}"] = sum(op({l: f.format(a=f"Z{c}", b=f"Z{c}") for l, f in fac.items()})
                                     for c in "rg")
        if self.dim_lat == 3:
            # theta-term corners E_nu C (SU2_operators.py:231-250): the first
            # corner of each plane times the electric field normal to it,
            # contracted with the spin-1/2 matrices over the rishon colours
            half = {"x": np.array([[0.0, 0.5], [0.5, 0.0]]), "y": np.array([[0.0, -0.5j], [0.5j, 0.0]]),
                    "z": np.array([[0.5, 0.0], [0.0, -0.5]])}
            planes = (("EzC_px,py", "z", {"px": "{1}_P", "py": "{2}_dag"}, 1j),
                      ("EyC_px,pz", "y", {"px": "{1}_P", "py": "P", "pz": "{2}_dag"}, -1j),
                      ("ExC_py,pz", "x", {"py": "{1}_P", "pz": "{2}_dag"}, 1j))
            for name, normal, fac, phase in planes:
                acc = 0
                for nu in "xyz":
                    E = ops[f"T{nu}_p{normal}"] + ops[f"T{nu}_m{normal}"]
                    for ii, c1 in enumerate("rg"):
                        for jj, c2 in enumerate("rg"):
                            f = half[nu][jj, ii]
                            if f == 0:
                                continue
                            corner = op({l: x.format("", f"Z{c1}", f"Z{c2}") for l, x in fac.items()})
                            acc = acc + phase * (corner @ (f * E))
                ops[name] = sp.csr_matrix(acc)
        if not self.pure:
            self._add_matter(ops, op)
        ops["E_square"] = sum(0.5 * ops[f"T2_{l}"] for l in links)
        self._add_background(ops, op)
        return ops

    def _generic_operators(self):
        """SU2_gen_dressed_site_operators (SU2_operators.py:302-592), any s_max, 1D-3D."""
        R = rishon_matrices_gen(self.s_max)
        Mt = None if self.pure else matter_matrices()
        links = self.links
        op = self._op_factory(R, Mt)
        ops = {}
        for name in ("T2", "P"):
            for l in links:
                ops[f"{name}_{l}"] = op({l: name})
        for l1, l2 in product("AB", repeat=2):
            for corner, fac in self._corner_lists().items():
                ops[f"C{l1}{l2}_{corner}"] = sum(
                    op({l: f.format(a=f"Z{l1}_{c}", b=f"Z{l2}_{c}") for l, f in fac.items()}) for c in "rg")
        if not self.pure:
            # Q1^dag = psi_r^dag Zg^dag - psi_g^dag Zr^dag, Q2^dag = psi_r^dag Zr + psi_g^dag Zg,
            # parities on the half-links before the acted one
            for k, l in enumerate(links):
                string = {links[q]: "P" for q in range(k)}
                ops[f"Q1_{l}_dag"] = (op({**string, l: "Zg_dag"}, matter="psi_r_dag_P")
                                      - op({**string, l: "Zr_dag"}, matter="psi_g_dag_P"))
                ops[f"Q2_{l}_dag"] = (op({**string, l: "Zr"}, matter="psi_r_dag_P")
                                      + op({**string, l: "Zg"}, matter="psi_g_dag_P"))
            for l in links:
                for q in ("Q1", "Q2"):
                    ops[f"{q}_{l}"] = ops[f"{q}_{l}_dag"].conj().T.tocsr()
            for lab in ("r", "g", "tot", "single", "pair", "zero"):
                ops[f"N_{lab}"] = op({}, matter=f"N_{lab}")
        ops["E_square"] = sum(0.5 * ops[f"T2_{l}"] for l in links)
        if self.bg > 0:
            ops["bg"] = op({}, background="T2")
        return ops

    def _add_matter(self, ops, op):
        """Hard-coded set: hoppings Q_{side axis}^dag = sum_c psi_c^dag Z_c with
        the parity string, number operators and N - 1."""
        links = self.links
        axes = "xyz"[: self.dim_lat]
        for a in axes:
            ops[f"Qm{a}_dag"] = 0
            ops[f"Qp{a}_dag"] = 0
        for c in "rg":
            # string of parities over the half-links before the acted one
            for k, l in enumerate(links):
                assign = {links[q]: "P" for q in range(k)}
                assign[l] = f"Z{c}"
                ops[f"Q{l}_dag"] = ops[f"Q{l}_dag"] + op(assign, matter=f"psi_{c}_dag_P")
        for a in axes:
            for s_ in "mp":
                ops[f"Q{s_}{a}"] = ops[f"Q{s_}{a}_dag"].conj().T.tocsr()
        for lab in ("r", "g", "tot", "single", "pair", "zero"):
            ops[f"N_{lab}"] = op({}, matter=f"N_{lab}")
        ops["N-1"] = ops["N_tot"] - sp.identity(self.full_dim, format="csr")

    def operators(self, label):
        """Every operator projected on the label's singlet basis, B^T O B."""
        if not hasattr(self, "_full"):
            self._full = self.full_operators()
        B = self.basis[label]
        return {k: np.asarray(B.T @ (v @ B)) for k, v in self._full.items()}
