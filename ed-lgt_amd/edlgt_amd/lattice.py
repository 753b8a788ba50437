"""Lattice geometry: site order, labels, bonds, plaquettes, link pairs.

Conventions of the reference (edlgt/modeling/lattice_mappings.py:11-94,
lattice_geometry.py:22-306, masks.py:18-42):
  * sites are numbered row-major with x fastest: i = x + Lx*y + Lx*Ly*z;
  * a bond along axis a from site i exists unless i sits on the upper OBC
    edge of a; the partner is the (periodic) next site along a;
  * a plaquette in plane (a, b) from site i is [i, i+a, i+b, i+a+b] and
    exists when both steps exist;
  * staggering: a site is even when (-1)^(x+y+z) > 0;
  * per-site basis labels are 'even'/'odd' ('site' without staggering),
    followed by '_mx'/'_px' ... for sites on lower/upper OBC edges.
"""
from math import prod

import numpy as np

AXES = "xyz"


class Lattice:
    def __init__(self, lvals, has_obc):
        self.lvals = [int(v) for v in lvals]
        self.has_obc = [bool(b) for b in has_obc]
        if len(self.lvals) != len(self.has_obc) or not 1 <= len(self.lvals) <= 3:
            raise ValueError("lvals and has_obc must have the same length 1..3")
        self.dim = len(self.lvals)
        self.n_sites = prod(self.lvals)
        self.axes = AXES[: self.dim]
        strides = [1]
        for L in self.lvals[:-1]:
            strides.append(strides[-1] * L)
        self._strides = strides

    def coords(self, i):
        return tuple((i // s) % L for s, L in zip(self._strides, self.lvals))

    def index(self, coords):
        return sum(c * s for c, s in zip(coords, self._strides))

    def step(self, i, axis):
        """Neighbour of site i along axis (periodic), or None at an OBC edge."""
        a = AXES.index(axis)
        c = list(self.coords(i))
        if c[a] == self.lvals[a] - 1 and self.has_obc[a]:
            return None
        c[a] = (c[a] + 1) % self.lvals[a]
        return self.index(c)

    def is_even(self, i):
        return sum(self.coords(i)) % 2 == 0

    def label(self, i, staggered):
        lab = ("even" if self.is_even(i) else "odd") if staggered else "site"
        for a, (c, L) in enumerate(zip(self.coords(i), self.lvals)):
            if self.has_obc[a]:
                if c == 0:
                    lab += f"_m{AXES[a]}"
                elif c == L - 1:
                    lab += f"_p{AXES[a]}"
        return lab

    def bonds(self, axis):
        """[(i, j)] in site order, j = i + axis."""
        out = []
        for i in range(self.n_sites):
            j = self.step(i, axis)
            if j is not None:
                out.append((i, j))
        return out

    def nbody_sites(self, i, distances):
        """NBodyTerm.get_nbody_neighbors (nbody_term.py:222-262): site i plus the
        sites at each displacement from it (periodic axes wrap), or None when a
        displacement leaves an open axis."""
        c0 = self.coords(i)
        out = [i]
        for dist in distances:
            c = list(c0)
            for a in range(self.dim):
                c[a] += int(dist[a])
                if not self.has_obc[a]:
                    c[a] %= self.lvals[a]
            if not all(0 <= c[a] < self.lvals[a] for a in range(self.dim)):
                return None
            out.append(self.index(c))
        return out

    def plaquettes(self, axes=("x", "y")):
        """[(i, i+a, i+b, i+a+b)] in site order."""
        out = []
        for i in range(self.n_sites):
            ia, ib = self.step(i, axes[0]), self.step(i, axes[1])
            if ia is None or ib is None:
                continue
            out.append((i, ia, ib, self.step(ia, axes[1])))
        return out

    def link_pairs(self):
        """(dir, site0, site1) for every link, direction-major, site order."""
        return [(d, i, j) for d, ax in enumerate(self.axes) for (i, j) in self.bonds(ax)]

    def origin_line(self, axis):
        """2D: sites with the other coordinate 0 along `axis` ('x': (x,0); 'y': (0,y))."""
        if self.dim != 2:
            raise NotImplementedError("origin lines are defined for 2D lattices")
        if axis == "x":
            return [self.index((x, 0)) for x in range(self.lvals[0])]
        return [self.index((0, y)) for y in range(self.lvals[1])]

    def origin_face(self, axis):
        """3D: sites of the coordinate plane through the origin orthogonal to
        `axis` ('x': x = 0, 'y': y = 0, 'z': z = 0), lattice_geometry.py:225-287."""
        if self.dim != 3:
            raise NotImplementedError("origin faces are defined for 3D lattices")
        a = AXES.index(axis)
        return [i for i in range(self.n_sites) if self.coords(i)[a] == 0]

    def even_mask(self):
        return np.array([self.is_even(i) for i in range(self.n_sites)])
