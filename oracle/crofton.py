"""Brute-force Crofton estimates of surface area, integral of mean curvature
and Euler characteristic of a digitised set (TEST INFRASTRUCTURE: only tests
import this module; the library never does).

The library's mkfCroftonMeasuresBPFD decomposes the Crofton formulas into
per-configuration weights of the 2x2x2 pattern histogram. This module
evaluates the same formulas globally, with no per-cell decomposition:

* direction weights: spherical Voronoi cells of the 26 lattice directions
  (scipy.spatial.SphericalVoronoi), a pair +-v getting twice one cell's area
  fraction;
* surface area S = 2 sum_v c_v N_v / |v|: N_v counts the lattice point pairs
  (x, x + v) inside the volume whose bits differ (Crofton: S_V = 2 P_L);
* integral of mean curvature M = 2 pi sum_v c_v d_v sum_planes chi: every
  lattice plane with normal v (spacing d_v) is cut out of the volume and the
  Euler number of the section is counted from connected components
  (scipy.ndimage.label): on square and rectangular plane lattices the mean of
  the 4- and 8-adjacency numbers, on the triangular lattices of the (1,1,1)
  planes the 6-adjacency number;
* Euler characteristic of the 6-adjacency foreground from the global counts
  of vertices, edges, faces and cubes of the cubical complex.

Method after Ohser & Muecklich, "Statistical Analysis of Microstructures in
Materials Science" (2000), ch. 4-5. The reference has no working counterpart
(its book measures need refMeasures.c, absent from its tree; src/mkfUtil.c:15-22).

The set must be padded by background on every face of the volume, so that
every boundary crossing and every section lies inside it.
"""
import itertools
import math

import numpy as np
from scipy import ndimage
from scipy.spatial import SphericalVoronoi


def direction_weights():
    """Area fraction of one direction's spherical Voronoi cell among the 26
    lattice directions, per class (1 axis, 2 face diagonal, 3 space diagonal)."""
    d = np.array([v for v in itertools.product((-1, 0, 1), repeat=3) if any(v)], dtype=float)
    u = d / np.linalg.norm(d, axis=1)[:, None]
    sv = SphericalVoronoi(u, radius=1.0, threshold=1e-12)
    sv.sort_vertices_of_regions()
    a = sv.calculate_areas() / (4 * math.pi)
    k = np.abs(d).sum(1)
    return {c: float(a[k == c].mean()) for c in (1, 2, 3)}


def _directions():
    """One vector of each +- pair of the 26 directions (13)."""
    out = []
    for v in itertools.product((-1, 0, 1), repeat=3):
        if any(v) and v > tuple(-c for c in v):
            out.append(v)
    return out


def _pairs_differ(b, v):
    """Number of pairs (p, p + v) inside the volume b[z, y, x] with b(p) != b(p + v)."""
    vx, vy, vz = v
    Z, Y, X = b.shape

    def sl(o, n):
        return slice(max(0, -o), n - max(0, o))

    src = b[sl(vz, Z), sl(vy, Y), sl(vx, X)]
    dst = b[sl(-vz, Z), sl(-vy, Y), sl(-vx, X)]
    return int(np.count_nonzero(src != dst))


def surface_area(b):
    w = direction_weights()
    s = 0.0
    for v in _directions():
        k = sum(abs(c) for c in v)
        s += 2 * w[k] * _pairs_differ(b, v) / math.sqrt(k)
    return 2 * s


_S4 = ndimage.generate_binary_structure(2, 1)
_S8 = ndimage.generate_binary_structure(2, 2)
_S6 = np.array([[0, 1, 1], [1, 1, 1], [1, 1, 0]], dtype=bool)  # offsets (+-1,0), (0,+-1), +-(1,-1)


def _chi2(a, fg, bg):
    p = np.pad(a, 1)
    nf = ndimage.label(p, structure=fg)[1]
    nb = ndimage.label(~p, structure=bg)[1]
    return nf - (nb - 1)


def _chi_square(a):
    return 0.5 * (_chi2(a, _S4, _S8) + _chi2(a, _S8, _S4))


def _axis_sections(b):
    """Sum of section Euler numbers over the planes normal to each axis."""
    tot = 0.0
    for ax in range(3):
        for k in range(b.shape[ax]):
            tot += _chi_square(np.take(b, k, axis=ax))
    return tot


def _diag_sections(b):
    """Planes x + y = c of b[z, y, x] (normal (1,1,0)): points (x, c - x, z),
    a rectangular lattice with steps (1,-1,0) and (0,0,1)."""
    Z, Y, X = b.shape
    tot = 0.0
    for c in range(X + Y - 1):
        xs = np.arange(max(0, c - Y + 1), min(X, c + 1))
        tot += _chi_square(b[:, c - xs, xs])
    return tot


def _tri_sections(b):
    """Planes x + y + z = c of b[z, y, x] (normal (1,1,1)): points
    (x, y, c - x - y) indexed by (x, y); triangular lattice, 6-adjacency."""
    Z, Y, X = b.shape
    xi, yi = np.meshgrid(np.arange(X), np.arange(Y), indexing="ij")
    tot = 0
    for c in range(X + Y + Z - 2):
        zi = c - xi - yi
        ok = (zi >= 0) & (zi < Z)
        sec = np.zeros((X, Y), dtype=bool)
        sec[ok] = b[zi[ok], yi[ok], xi[ok]]
        tot += _chi2(sec, _S6, _S6)
    return tot


def mean_curvature(b):
    w = direction_weights()
    b = np.asarray(b, dtype=bool)
    m = 2 * w[1] * _axis_sections(b)
    # the six face-diagonal normals: (1,1,0) in each axis pair, and its mirror
    face = 0.0
    for t in ((0, 1, 2), (1, 0, 2), (2, 1, 0)):  # bring each axis pair to (y, x)
        bt = np.transpose(b, t)
        face += _diag_sections(bt) + _diag_sections(bt[:, ::-1, :])
    m += 2 * w[2] * face / math.sqrt(2)
    space = 0.0
    for flip in (None, 0, 1, 2):
        bf = b if flip is None else np.flip(b, axis=flip)
        space += _tri_sections(bf)
    m += 2 * w[3] * space / math.sqrt(3)
    return 2 * math.pi * m


def euler_6(b):
    """Euler characteristic of the cubical complex of set vertices, edges
    between set axis neighbours, fully set faces and cubes (6-adjacency)."""
    b = np.asarray(b, dtype=bool)
    v = int(b.sum())
    e = sum(int(np.count_nonzero(b[tuple(slice(None, -1) if i == a else slice(None) for i in range(3))] &
                                 b[tuple(slice(1, None) if i == a else slice(None) for i in range(3))]))
            for a in range(3))
    f = 0
    for a in range(3):
        s = [i for i in range(3) if i != a]
        q = None
        for d0 in (0, 1):
            for d1 in (0, 1):
                idx = [slice(None)] * 3
                idx[s[0]] = slice(d0, b.shape[s[0]] - 1 + d0)
                idx[s[1]] = slice(d1, b.shape[s[1]] - 1 + d1)
                t = b[tuple(idx)]
                q = t if q is None else q & t
        f += int(np.count_nonzero(q))
    c = np.ones(tuple(n - 1 for n in b.shape), dtype=bool)
    for dz, dy, dx in itertools.product((0, 1), repeat=3):
        c &= b[dz:dz + b.shape[0] - 1, dy:dy + b.shape[1] - 1, dx:dx + b.shape[2] - 1]
    return v - e + f - int(np.count_nonzero(c))
