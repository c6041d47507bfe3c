"""Reference tetrahedron for continuous-Galerkin P1/P2 elements.

The reference element is the tetrahedron with vertices (-1,-1,-1), (1,-1,-1),
(-1,1,-1), (-1,-1,1) in (r,s,t) coordinates, as set up by Build_reference
(src/mod_geometry.f90:2307-2523).  Nodes are the equidistant lattice points in
the order i (along r) fastest, then j (s), then k (t); for P2 this gives the
local order ``vertices 1,3,6,10; edge nodes 2,4,5,7,8,9 on edges (1-2), (1-3),
(2-3), (1-4), (2-4), (3-4)`` (src/mod_geometry.f90:1575-1586, 1707-1711).
For pOrder <= 2 the warp-and-blend node set of the reference coincides with the
equidistant one.

The mass matrix M_ij = int l_i l_j and the differentiation matrices
Drst[d]_ij = dl_j/dr_d (node_i) are mathematically those of the reference
(MassM = Vinv^T Vinv and Drst = GradV Vinv with an orthonormal Dubiner basis,
src/mod_geometry.f90:2459-2487).  Here they are computed exactly in rational
arithmetic from a monomial basis on the unit simplex and rounded once to
float64.
"""
from __future__ import annotations

from fractions import Fraction
from functools import lru_cache
from math import factorial

import numpy as np

__all__ = ["reference_element", "vertex_slots", "n_nodes"]


def n_nodes(porder: int) -> int:
    return (porder + 1) * (porder + 2) * (porder + 3) // 6


def vertex_slots(porder: int) -> tuple[int, int, int, int]:
    """0-based local slots of the four vertices (reference ``vord``)."""
    if porder == 1:
        return (0, 1, 2, 3)
    if porder == 2:
        return (0, 2, 5, 9)
    raise ValueError("pOrder must be 1 or 2")


def _lattice(porder: int):
    pts = []
    for k in range(porder + 1):
        for j in range(porder + 1 - k):
            for i in range(porder + 1 - k - j):
                pts.append((Fraction(i, porder), Fraction(j, porder), Fraction(k, porder)))
    return pts  # barycentric-like unit simplex coordinates (xi, eta, zeta)


def _monomials(porder: int):
    return [(a, b, c) for a in range(porder + 1) for b in range(porder + 1 - a)
            for c in range(porder + 1 - a - b)]


def _solve_exact(V, rhs_cols):
    """Solve V X = R exactly (Gauss-Jordan over Fractions)."""
    n = len(V)
    A = [list(V[i]) + list(rhs_cols[i]) for i in range(n)]
    m = len(A[0])
    for c in range(n):
        p = next(r for r in range(c, n) if A[r][c] != 0)
        A[c], A[p] = A[p], A[c]
        piv = A[c][c]
        A[c] = [x / piv for x in A[c]]
        for r in range(n):
            if r != c and A[r][c] != 0:
                f = A[r][c]
                A[r] = [A[r][k] - f * A[c][k] for k in range(m)]
    return [row[n:] for row in A]


@lru_cache(maxsize=4)
def reference_element(porder: int):
    """Return (nodes_rst[pNp,3], MassM[pNp,pNp], Drst[3,pNp,pNp]) as float64."""
    if porder not in (1, 2):
        raise ValueError("pOrder must be 1 or 2")
    lat = _lattice(porder)
    mons = _monomials(porder)
    n = len(lat)
    # Vandermonde V[i][a] = m_a(node_i); nodal basis coefficients C = V^{-1}
    V = [[x ** a * y ** b * z ** c for (a, b, c) in mons] for (x, y, z) in lat]
    eye = [[Fraction(int(i == j)) for j in range(n)] for i in range(n)]
    C = _solve_exact(V, eye)  # C[a][j]: coefficient of monomial a in l_j

    def simplex_int(a, b, c):
        # int over the unit simplex of xi^a eta^b zeta^c; the reference tet is
        # the image under r = 2 xi - 1, so volumes scale by 8.
        return Fraction(factorial(a) * factorial(b) * factorial(c), factorial(a + b + c + 3))

    G = [[8 * simplex_int(ma[0] + mb[0], ma[1] + mb[1], ma[2] + mb[2]) for mb in mons] for ma in mons]
    M = [[sum(C[a][i] * G[a][b] * C[b][j] for a in range(n) for b in range(n))
          for j in range(n)] for i in range(n)]

    def dmon(mon, d, pt):
        e = list(mon)
        if e[d] == 0:
            return Fraction(0)
        coef = e[d]
        e[d] -= 1
        return coef * pt[0] ** e[0] * pt[1] ** e[1] * pt[2] ** e[2]

    D = []
    for d in range(3):
        # d/dr_d = (1/2) d/dxi_d
        D.append([[Fraction(1, 2) * sum(dmon(mons[a], d, lat[i]) * C[a][j] for a in range(n))
                   for j in range(n)] for i in range(n)])

    nodes = np.array([[float(2 * x - 1) for x in p] for p in lat])
    Mf = np.array([[float(v) for v in row] for row in M])
    Df = np.array([[[float(v) for v in row] for row in Dd] for Dd in D])
    return nodes, Mf, Df


@lru_cache(maxsize=4)
def reference_faces(porder: int):
    """Face node lists and face mass matrices of the reference tetrahedron.

    Face f (0-based) is opposite vertex f: r+s+t = -1, r = -1, s = -1, t = -1
    (src/mod_geometry.f90:2331-2347, 2389-2409).  ``fmask[f]`` lists its nodes
    in increasing local order; ``massF[f]`` = int phi_a phi_b over the
    reference triangle (-1,-1), (1,-1), (-1,1) in the face coordinates the
    reference projects on ((s,t), (s,t), (r,t), (r,s); :2495-2513, where it is
    formed as inv(V2D V2D^T) of an orthonormal basis).  Computed exactly in
    rational arithmetic and rounded once.  Returns (fmask[4, nfp] int32,
    massF[4, nfp, nfp] float64)."""
    if porder not in (1, 2):
        raise ValueError("pOrder must be 1 or 2")
    lat = _lattice(porder)                      # (xi, eta, zeta) = (r+1, s+1, t+1)/2
    tests = (lambda p: p[0] + p[1] + p[2] == 1, lambda p: p[0] == 0, lambda p: p[1] == 0, lambda p: p[2] == 0)
    coords = ((1, 2), (1, 2), (0, 2), (0, 1))
    mons = [(a, b) for a in range(porder + 1) for b in range(porder + 1 - a)]

    def tri_int(a, b):
        # int over the unit triangle of xi^a eta^b; the reference triangle has 4x the area
        return Fraction(factorial(a) * factorial(b), factorial(a + b + 2))

    G = [[4 * tri_int(ma[0] + mb[0], ma[1] + mb[1]) for mb in mons] for ma in mons]
    fmask, mass = [], []
    for f in range(4):
        idx = [i for i, p in enumerate(lat) if tests[f](p)]
        pts = [(lat[i][coords[f][0]], lat[i][coords[f][1]]) for i in idx]
        n = len(pts)
        V = [[x ** a * y ** b for (a, b) in mons] for (x, y) in pts]
        eye = [[Fraction(int(i == j)) for j in range(n)] for i in range(n)]
        C = _solve_exact(V, eye)
        M = [[sum(C[a][i] * G[a][b] * C[b][j] for a in range(n) for b in range(n)) for j in range(n)]
             for i in range(n)]
        fmask.append(idx)
        mass.append([[float(v) for v in row] for row in M])
    return np.array(fmask, dtype=np.int32), np.array(mass, dtype=np.float64)
