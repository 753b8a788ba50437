"""Lattice geometry of the reference (oracle restatement; test infrastructure).

Follows edlgt/modeling/lattice_mappings.py:11-94 (row-major, x fastest),
edlgt/modeling/lattice_geometry.py:22-306 (labels, neighbours, plaquettes,
origin surfaces, link pairs) and edlgt/modeling/masks.py:18-42 (staggering).
"""
from itertools import product
from math import prod

import numpy as np


def zig_zag(lvals, idx):
    """Linear site index -> coords, x fastest (lattice_mappings.py:20-53)."""
    coords = []
    stride = 1
    for L in lvals:
        coords.append((idx // stride) % L)
        stride *= L
    return tuple(coords)


def inverse_zig_zag(lvals, coords):
    """coords -> linear index (lattice_mappings.py:56-94)."""
    idx, stride = 0, 1
    for c, L in zip(coords, lvals):
        idx += c * stride
        stride *= L
    return idx


def site_label(coords, lvals, has_obc, staggered):
    """Per-site basis label, e.g. 'even_mx_py' (lattice_geometry.py:22-74)."""
    dims = "xyz"
    if staggered:
        stag = "even" if (-1) ** sum(coords) > 0 else "odd"
    else:
        stag = "site"
    lab = stag
    for ii, c in enumerate(coords):
        if has_obc[ii]:
            if c == 0:
                lab += f"_m{dims[ii]}"
            elif c == lvals[ii] - 1:
                lab += f"_p{dims[ii]}"
    return lab


def neighbor_sites(coords, lvals, axis, has_obc):
    """Bond [i, i+axis] or None at an OBC edge (lattice_geometry.py:110-163)."""
    ax = "xyz".index(axis)
    if coords[ax] < lvals[ax] - 1 or not has_obc[ax]:
        c2 = list(coords)
        c2[ax] = (c2[ax] + 1) % lvals[ax]
        return [inverse_zig_zag(lvals, coords), inverse_zig_zag(lvals, c2)]
    return None


def plaquette_sites(coords, lvals, axes, has_obc):
    """[origin, +a, +b, +a+b] or None (lattice_geometry.py:166-222)."""
    ia, ib = "xyz".index(axes[0]), "xyz".index(axes[1])
    for i in (ia, ib):
        if not (coords[i] < lvals[i] - 1 or not has_obc[i]):
            return None
    base = list(coords)
    c1 = list(base); c1[ia] = (c1[ia] + 1) % lvals[ia]
    c2 = list(base); c2[ib] = (c2[ib] + 1) % lvals[ib]
    c3 = list(base); c3[ia] = (c3[ia] + 1) % lvals[ia]; c3[ib] = (c3[ib] + 1) % lvals[ib]
    return [inverse_zig_zag(lvals, c) for c in (base, c1, c2, c3)]


def link_pairs(lvals, has_obc):
    """Per direction, the (site_i, site_i+dir) pairs in row-major order
    (lattice_geometry.py:290-306)."""
    out = []
    for d in "xyz"[: len(lvals)]:
        lst = []
        for ii in range(prod(lvals)):
            s = neighbor_sites(zig_zag(lvals, ii), lvals, d, has_obc)
            if s is not None:
                lst.append(s)
        out.append(np.array(lst, dtype=np.int64).reshape(-1, 2))
    return out


def origin_lines_2d(lvals):
    """2D origin edges: 'x' = sites (x,0), 'y' = sites (0,y) (lattice_geometry.py:225-287)."""
    Lx, Ly = lvals
    return {
        "x": [inverse_zig_zag(lvals, (x, 0)) for x in range(Lx)],
        "y": [inverse_zig_zag(lvals, (0, y)) for y in range(Ly)],
    }


def staggered_mask(lvals, label):
    """Boolean mask of shape lvals: even <=> (-1)^sum(coords) > 0 (masks.py:18-42)."""
    m = np.zeros(lvals, dtype=bool)
    for c in product(*[range(L) for L in lvals]):
        s = (-1) ** sum(c)
        m[c] = (label == "even" and s > 0) or (label == "odd" and s < 0)
    return m


def border_labels(lattice_dim):
    """lattice_geometry.py:309-342: every non-empty set of axes (in x, y, z
    order) with a lower (m) or upper (p) side per axis, e.g. 'px_my'."""
    from itertools import combinations, product
    axes = "xyz"[:lattice_dim]
    out = []
    for k in range(1, lattice_dim + 1):
        for combo in combinations(axes, k):
            for sides in product("mp", repeat=k):
                out.append("_".join(f"{sd}{a}" for sd, a in zip(sides, combo)))
    return out


def border_configs(links, lattice_dim, offset=0):
    """Which border subsets a local link configuration belongs to
    (LGT_border_configs, lattice_geometry.py:345-487): a label applies when
    every half-link it names holds the offset value; links are ordered
    (m_x, m_y, ..., p_x, p_y, ...)."""
    pos = {f"m{a}": i for i, a in enumerate("xyz"[:lattice_dim])}
    pos.update({f"p{a}": lattice_dim + i for i, a in enumerate("xyz"[:lattice_dim])})
    return [lab for lab in border_labels(lattice_dim)
            if all(links[pos[part]] == offset for part in lab.split("_"))]


def origin_faces_3d(lvals):
    """Coordinate planes through the origin (lattice_geometry.py:225-287):
    'yz' (x=0), 'xz' (y=0), 'xy' (z=0) as site lists."""
    Lx, Ly, Lz = lvals
    return {"yz": [inverse_zig_zag(lvals, (0, y, z)) for y in range(Ly) for z in range(Lz)],
            "xz": [inverse_zig_zag(lvals, (x, 0, z)) for x in range(Lx) for z in range(Lz)],
            "xy": [inverse_zig_zag(lvals, (x, y, 0)) for x in range(Lx) for y in range(Ly)]}
