"""Seeded tetrahedral meshes of a ball and their P1/P2 node sets.

The reference reads TetGen meshes produced by an external builder
(PlanetaryModels; README.md:41) through ``pnm_read_ele``/``pnm_read_node``
(src/mod_geometry.f90:140-425) and adds P2 edge nodes in ``pnm_add_edges``
(src/mod_geometry.f90:428-689).  This module replaces that producer with a
deterministic generator: a cubic lattice of spacing h with uniform jitter
+-jitter*h, points on the bounding sphere, Delaunay tetrahedralisation (Qhull),
removal of flat hull slivers, positive orientation, P2 edge nodes at edge
midpoints and a Morton (Z-order) node numbering for memory locality of the
assembled CSR rows.  ``stuffed_ball_mesh`` builds the layered bodies without
slivers instead: the body-centred cubic lattice cut by each interface sphere
with isosurface stuffing.

Element-node layout follows the reference (see ``refelem``): for P2 the local
slots are vertices (0,2,5,9) and edges (1-2)->1, (1-3)->3, (2-3)->4, (1-4)->6,
(2-4)->7, (3-4)->8 (0-based slots, 1-based vertex names).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["TetMesh", "ball_mesh", "ellipsoid_mesh", "element_radius", "h_for_tets", "mesh_from_vertices",
           "stuffed_ball_mesh"]

# (vertex a, vertex b, slot) for P2 edge nodes; vertices are 0-based here
_P2_EDGES = ((0, 1, 1), (0, 2, 3), (1, 2, 4), (0, 3, 6), (1, 3, 7), (2, 3, 8))
_P2_VERT_SLOTS = (0, 2, 5, 9)


@dataclass
class TetMesh:
    nodes: np.ndarray        # (n_nodes, 3) float64, all FE nodes (P1: vertices; P2: + edge midpoints)
    elems: np.ndarray        # (n_elem, pNp) int32 0-based node ids in reference local order
    tets: np.ndarray         # (n_elem, 4) int32 vertex node ids
    porder: int
    surface: np.ndarray      # (n_nodes,) bool: node lies on the outer surface
    seed: int
    h: float
    vlist: np.ndarray | None = None  # (n_nodes,) input node id of each node (meshes read from files)

    @property
    def n_nodes(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def n_elem(self) -> int:
        return int(self.elems.shape[0])

    @property
    def n_dof(self) -> int:
        return 3 * self.n_nodes


def _fibonacci_sphere(n: int, rng) -> np.ndarray:
    i = np.arange(n) + 0.5
    phi = np.arccos(1.0 - 2.0 * i / n)
    theta = np.pi * (1.0 + 5.0 ** 0.5) * i + rng.uniform(0, 2 * np.pi)
    return np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], axis=1)


def element_radius(X: np.ndarray, T: np.ndarray) -> np.ndarray:
    """The radius that places each tetrahedron in a layer: the mean of its
    vertex radii, or, for a tetrahedron with all four vertices on one sphere
    (a chord polytope, inside it), its centroid's radius.  A tetrahedron with
    a face on a sphere and its fourth vertex just outside can have its
    centroid inside; the mean vertex radius keeps it outside."""
    rv = np.linalg.norm(X[T], axis=2)
    rm = rv.mean(axis=1)
    flat = (rv.max(axis=1) - rv.min(axis=1)) <= 1e-12 * np.maximum(rv.max(axis=1), 1e-300)
    return np.where(flat, np.linalg.norm(X[T].mean(axis=1), axis=1), rm)


def _tet_volumes(X: np.ndarray, T: np.ndarray) -> np.ndarray:
    a = X[T[:, 1]] - X[T[:, 0]]
    b = X[T[:, 2]] - X[T[:, 0]]
    c = X[T[:, 3]] - X[T[:, 0]]
    return np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0


def _morton_order(X: np.ndarray) -> np.ndarray:
    lo = X.min(axis=0)
    span = float((X.max(axis=0) - lo).max()) or 1.0
    q = np.clip(((X - lo) / span * (2 ** 20 - 1)).astype(np.uint64), 0, 2 ** 20 - 1)

    def spread(v):
        v = v & np.uint64(0x1FFFFF)
        v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
        v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
        v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
        v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
        return v

    code = spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1)) | (spread(q[:, 2]) << np.uint64(2))
    return np.argsort(code, kind="stable")


def h_for_tets(n_tets: int, radius: float = 1.0) -> float:
    """Lattice spacing giving about ``n_tets`` tetrahedra in a ball (empirical 6.74 tets per lattice cell)."""
    vol = 4.0 / 3.0 * np.pi * radius ** 3
    return float((vol * 6.74 / n_tets) ** (1.0 / 3.0))


# interface spheres: sampled at INTERFACE_SPACING * h, lattice points kept at
# least INTERFACE_BAND * h away (measured: with these values Delaunay keeps
# every interface triangle as a tetrahedron face from ~1k tets on)
INTERFACE_SPACING = 0.6
INTERFACE_BAND = 0.7


def ball_mesh(n_tets: int | None = None, *, h: float | None = None, porder: int = 2, seed: int = 1019,
              jitter: float = 0.25, radius: float = 1.0, axes=(1.0, 1.0, 1.0), interfaces=(),
              scale: float = 1.0) -> TetMesh:
    """Jittered-lattice Delaunay mesh of a ball (or, with ``axes``, an ellipsoid).

    ``interfaces``: radii (fractions of ``radius``) of internal spheres, such
    as the boundaries of a fluid outer core.  Points are placed on each
    sphere and kept off a band around it, so that every tetrahedron lies on
    one side: the interface is a surface of tetrahedron faces, as in the
    reference's input meshes.  A mesh that fails this check raises.
    ``scale`` multiplies the coordinates at the end (e.g. 6371 for a body in
    km with the layer radii given as fractions)."""
    from scipy.spatial import Delaunay

    if h is None:
        if n_tets is None:
            raise ValueError("give n_tets or h")
        h = h_for_tets(n_tets, radius)
    rng = np.random.default_rng(seed)
    n = int(np.ceil(radius / h)) + 1
    g = np.arange(-n, n + 1, dtype=np.float64) * h
    P = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    P = P + rng.uniform(-jitter * h, jitter * h, size=P.shape)
    rP = np.linalg.norm(P, axis=1)
    keep = rP < radius - 0.55 * h
    radii = [float(f) * radius for f in interfaces]
    for R in radii:
        keep &= np.abs(rP - R) > INTERFACE_BAND * h
    P = P[keep]
    parts = [P]
    for R in radii:
        hs = INTERFACE_SPACING * h
        parts.append(R * _fibonacci_sphere(max(12, int(round(4.0 * np.pi * R * R / (hs * hs * np.sqrt(3.0) / 2.0)))),
                                           rng))
    ns = max(12, int(round(4.0 * np.pi * radius ** 2 / (h * h * np.sqrt(3.0) / 2.0))))
    S = radius * _fibonacci_sphere(ns, rng)
    X = np.concatenate(parts + [S], axis=0)
    on_surf = np.zeros(X.shape[0], dtype=bool)
    on_surf[X.shape[0] - S.shape[0]:] = True

    T = Delaunay(X, qhull_options="Qbb Qc Qz Q12").simplices.astype(np.int64)
    vol = _tet_volumes(X, T)
    flip = vol < 0
    T[flip, 2], T[flip, 3] = T[flip, 3].copy(), T[flip, 2].copy()
    vol = np.abs(vol)
    # flat slivers whose four vertices all lie on the sphere are hull artefacts
    all_surf = on_surf[T].all(axis=1)
    edges = np.stack([X[T[:, b]] - X[T[:, a]] for a, b, _ in _P2_EDGES], axis=1)
    lrms = np.sqrt((edges ** 2).sum(axis=2).mean(axis=1))
    quality = 6.0 * np.sqrt(2.0) * vol / lrms ** 3
    keep = ~(all_surf & (quality < 0.15)) & (vol > 1e-12 * h ** 3)
    T = T[keep]
    used = np.unique(T)
    remap = -np.ones(X.shape[0], dtype=np.int64)
    remap[used] = np.arange(used.size)
    X = X[used]
    on_surf = on_surf[used]
    T = remap[T]
    if radii:
        _check_conforming(X, T, radii)

    ax = np.asarray(axes, dtype=np.float64) * float(scale)
    X = X * ax[None, :]
    return _assemble_mesh(X, T, on_surf, porder, seed=seed, h=h * float(scale), reorder_elems=True)


# Isosurface stuffing (F. Labelle and J. R. Shewchuk, "Isosurface stuffing:
# fast tetrahedral meshes with good dihedral angles", SIGGRAPH 2007): warp
# thresholds for the BCC lattice's axis-aligned edges (length h) and its
# diagonal edges (length sqrt(3)/2 h).  A cut point closer than alpha * length
# to an edge end moves that end onto the surface.
STUFF_ALPHA_LONG = 0.24999
STUFF_ALPHA_SHORT = 0.41189
BCC_TETS_PER_CELL = 12


def stuffed_ball_mesh(n_tets: int | None = None, *, h: float | None = None, porder: int = 2, seed: int = 1019,
                      radius: float = 1.0, interfaces=(), scale: float = 1.0) -> TetMesh:
    """Sliver-free mesh of a ball with internal interface spheres: the
    body-centred cubic lattice (twelve congruent tetrahedra per cell, dihedral
    angles of 60 and 90 degrees), shifted by a seeded offset, cut by each
    sphere with isosurface stuffing (lattice vertices near a sphere warped
    onto it, crossing tetrahedra split at the cut points), both sides kept at
    an interface and the outside dropped at the surface.

    The Delaunay meshes of ``ball_mesh`` hold slivers (quality ~1e-3 at 1e5
    tetrahedra) that raise the pencil's largest eigenvalue by three to four
    orders of magnitude, and with it every Chebyshev degree; these meshes keep
    the largest eigenvalue at the size of the lattice spacing.  Interfaces
    are surfaces of tetrahedron faces, as in ``ball_mesh``."""
    if h is None:
        if n_tets is None:
            raise ValueError("give n_tets or h")
        h = float((4.0 / 3.0 * np.pi * radius ** 3 * BCC_TETS_PER_CELL / n_tets) ** (1.0 / 3.0))
    rng = np.random.default_rng(seed)
    off = rng.uniform(0.0, h, size=3)
    X, T = _bcc_lattice(h, radius, off)
    radii = sorted(float(f) * radius for f in interfaces)
    for R in radii:
        X, T = _stuff_sphere(X, T, R, h, keep_outside=True)
        X, T = _split_chords(X, T, R)
    X, T = _stuff_sphere(X, T, radius, h, keep_outside=False)
    used = np.unique(T)
    remap = -np.ones(X.shape[0], dtype=np.int64)
    remap[used] = np.arange(used.size)
    X = X[used]
    T = remap[T]
    rX = np.linalg.norm(X, axis=1)
    on_surf = np.abs(rX - radius) <= 1e-12 * radius
    if radii:
        _check_conforming(X, T, radii)
    return _assemble_mesh(X * float(scale), T, on_surf, porder, seed=seed, h=h * float(scale),
                          reorder_elems=True)


def _bcc_lattice(h, radius, off):
    """Vertices and tetrahedra of the BCC lattice over the cube [-R-2h, R+2h]^3
    (shifted by ``off``); only tetrahedra with a vertex within the ball plus
    one spacing are kept.  Each pair of face-adjacent cell centres spans four
    tetrahedra with the four edges of their shared face."""
    n = int(np.ceil(radius / h)) + 2
    g = np.arange(-n, n + 1)
    nc = g.size
    corners = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3) * h + off
    gc = np.arange(-n, n)
    ncc = gc.size
    centres = (np.stack(np.meshgrid(gc, gc, gc, indexing="ij"), axis=-1).reshape(-1, 3) + 0.5) * h + off
    X = np.concatenate([corners, centres], axis=0)

    def cid(i, j, k):
        return ((i + n) * nc + (j + n)) * nc + (k + n)

    def zid(i, j, k):
        return corners.shape[0] + ((i + n) * ncc + (j + n)) * ncc + (k + n)

    I, J, K = (a.ravel() for a in np.meshgrid(gc, gc, gc, indexing="ij"))
    tets = []
    for d in range(3):
        ok = [I, J, K][d] < n - 1
        i, j, k = I[ok], J[ok], K[ok]
        c0 = zid(i, j, k)
        nb = [i, j, k]
        nb[d] = nb[d] + 1
        c1 = zid(*nb)
        # the shared face: coordinate d fixed at cell index + 1, the other two
        # axes (e, f) over {0, 1}; its four edges in cyclic order
        e, f = [a for a in range(3) if a != d]
        quad = []
        for de, df in ((0, 0), (1, 0), (1, 1), (0, 1)):
            v = [i, j, k]
            v = [a.copy() for a in v]
            v[d] = v[d] + 1
            v[e] = v[e] + de
            v[f] = v[f] + df
            quad.append(cid(*v))
        for q in range(4):
            tets.append(np.stack([c0, c1, quad[q], quad[(q + 1) % 4]], axis=1))
    T = np.concatenate(tets, axis=0)
    near = (np.linalg.norm(X[T], axis=2).min(axis=1) < radius + h)
    T = T[near]
    return X, _orient(X, T)


def _orient(X, T):
    vol = _tet_volumes(X, T)
    T = T.copy()
    flip = vol < 0
    T[flip, 2], T[flip, 3] = T[flip, 3].copy(), T[flip, 2].copy()
    return T


_TET_EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))


def _sphere_cut(Xa, Xb, R):
    """Parameter t in (0, 1) where the segment Xa + t (Xb - Xa) meets |x| = R
    (one end inside, one outside: exactly one root in the interval)."""
    d = Xb - Xa
    A = np.einsum("ij,ij->i", d, d)
    B = 2.0 * np.einsum("ij,ij->i", Xa, d)
    C = np.einsum("ij,ij->i", Xa, Xa) - R * R
    disc = np.sqrt(np.maximum(B * B - 4.0 * A * C, 0.0))
    t1 = (-B + disc) / (2.0 * A)
    t2 = (-B - disc) / (2.0 * A)
    return np.where((t1 >= 0.0) & (t1 <= 1.0), t1, t2)


def _stuff_sphere(X, T, R, h, keep_outside):
    """One sphere of isosurface stuffing over the tetrahedra T: warp, then cut."""
    X = X.copy()
    r = np.linalg.norm(X, axis=1)
    phi = r - R
    on = np.abs(phi) <= 1e-12 * R
    X[on] *= (R / r[on])[:, None]
    phi[on] = 0.0
    nv = X.shape[0]
    pairs = np.unique(np.sort(np.concatenate([T[:, [a, b]] for a, b in _TET_EDGES], axis=0), axis=1), axis=0)
    a, b = pairs[:, 0], pairs[:, 1]
    cross = phi[a] * phi[b] < 0.0
    a, b = a[cross], b[cross]
    t = _sphere_cut(X[a], X[b], R)
    L = np.linalg.norm(X[b] - X[a], axis=1)
    alpha = np.where(L > 0.93 * h, STUFF_ALPHA_LONG, STUFF_ALPHA_SHORT)
    P = X[a] + t[:, None] * (X[b] - X[a])
    # each violated end moves to its closest violating cut point
    best = np.full(nv, np.inf)
    target = np.zeros((nv, 3))
    for end, dist, viol in ((a, t * L, t < alpha), (b, (1.0 - t) * L, (1.0 - t) < alpha)):
        e, dd, pp = end[viol], dist[viol], P[viol]
        order = np.argsort(-dd, kind="stable")     # the closest written last
        e, dd, pp = e[order], dd[order], pp[order]
        closer = dd < best[e]
        best[e[closer]] = dd[closer]
        target[e[closer]] = pp[closer]
    moved = np.isfinite(best)
    X[moved] = target[moved]
    phi[moved] = 0.0
    # the closest-first write above may leave a farther point where two cut
    # points tie in one scatter: recheck that every moved vertex is on the sphere
    X[moved] *= (R / np.linalg.norm(X[moved], axis=1))[:, None]

    s = np.sign(phi[T])
    crossing = (s < 0).any(axis=1) & (s > 0).any(axis=1)
    whole = T[~crossing]
    ws = s[~crossing]
    inside = (ws <= 0).all(axis=1)
    keep = [whole[inside]]
    if keep_outside:
        keep.append(whole[~inside])
    # cut points of the crossing tetrahedra, one new vertex per cut edge
    Tc = T[crossing]
    cp = np.concatenate([np.sort(Tc[:, [p, q]], axis=1) for p, q in _TET_EDGES], axis=0)
    cp = cp[phi[cp[:, 0]] * phi[cp[:, 1]] < 0.0]
    cp = np.unique(cp, axis=0)
    tcut = _sphere_cut(X[cp[:, 0]], X[cp[:, 1]], R)
    Pc = X[cp[:, 0]] + tcut[:, None] * (X[cp[:, 1]] - X[cp[:, 0]])
    Pc *= (R / np.linalg.norm(Pc, axis=1))[:, None]
    cut_id = {(int(p), int(q)): nv + k for k, (p, q) in enumerate(cp)}
    X = np.concatenate([X, Pc], axis=0)
    pieces_in, pieces_out = [], []
    for tet in Tc:
        sg = np.sign(phi[tet])
        neg = [int(v) for v, sv in zip(tet, sg) if sv < 0]
        pos = [int(v) for v, sv in zip(tet, sg) if sv > 0]
        zer = [int(v) for v, sv in zip(tet, sg) if sv == 0]

        def c(u, w):
            return cut_id[(u, w) if u < w else (w, u)]

        for side, own, other, out in ((-1, neg, pos, pieces_in), (1, pos, neg, pieces_out)):
            if side > 0 and not keep_outside:
                continue
            out.extend(_piece(own, zer, other, c))
    keep.append(np.asarray(pieces_in, dtype=np.int64).reshape(-1, 4))
    if keep_outside:
        keep.append(np.asarray(pieces_out, dtype=np.int64).reshape(-1, 4))
    T = np.concatenate(keep, axis=0)
    T = _orient(X, T)
    vol = _tet_volumes(X, T)
    if np.any(vol <= 1e-9 * h ** 3):
        raise ValueError("isosurface stuffing produced a degenerate tetrahedron")
    return X, T


def _piece(own, zer, other, c):
    """Tetrahedra of one side of a cut tetrahedron: the convex hull of the
    side's own vertices, the vertices on the sphere and the cut points of the
    edges between own and other vertices.  Quadrilateral faces are split
    through their smallest vertex id, so that neighbours agree."""
    no, nz = len(own), len(zer)
    if no == 1 and nz == 0:                       # tetrahedron at a vertex
        v = own[0]
        return [[v] + [c(v, w) for w in other]]
    if no == 3 and nz == 0:                       # prism: own triangle, cut triangle
        return _prism(own, [c(v, other[0]) for v in own])
    if no == 2 and nz == 0:                       # prism between the two own vertices
        m1, m2 = own
        p1, p2 = other
        return _prism([m1, c(m1, p1), c(m1, p2)], [m2, c(m2, p1), c(m2, p2)])
    if no == 1 and nz == 1:                       # tetrahedron
        v = own[0]
        return [[v, zer[0]] + [c(v, w) for w in other]]
    if no == 2 and nz == 1:                       # pyramid: apex on the sphere
        p1, p2 = own
        m = other[0]
        return _pyramid(zer[0], [p1, p2, c(p2, m), c(p1, m)])
    if no == 1 and nz == 2:                       # tetrahedron
        return [[own[0], zer[0], zer[1], c(own[0], other[0])]]
    raise ValueError(f"unexpected cut pattern own={no} zero={nz}")


def _pyramid(apex, q):
    i = int(np.argmin(q))
    if i in (0, 2):
        return [[apex, q[0], q[1], q[2]], [apex, q[0], q[2], q[3]]]
    return [[apex, q[1], q[2], q[3]], [apex, q[1], q[3], q[0]]]


def _prism(bot, top):
    """Three tetrahedra of the prism bot[i] -- top[i], every quadrilateral
    face split through its smallest vertex id (always realisable: the
    smallest vertex of the prism sees both of its quadrilaterals' diagonals)."""
    allv = list(bot) + list(top)
    k = int(np.argmin(allv))
    if k >= 3:
        bot, top = top, bot
        k -= 3
    a0, a1, a2 = bot[k], bot[(k + 1) % 3], bot[(k + 2) % 3]
    b0, b1, b2 = top[k], top[(k + 1) % 3], top[(k + 2) % 3]
    if min(a1, b2) < min(a2, b1):                 # diagonal a1 - b2
        return [[a0, a1, a2, b2], [a0, a1, b2, b1], [a0, b1, b2, b0]]
    return [[a0, a1, a2, b1], [a0, a2, b2, b1], [a0, b1, b2, b0]]


def _split_chords(X, T, R):
    """Bisect every edge between two vertices on the interface sphere r = R
    whose tetrahedra all lie on one side of it (a chord through one region):
    the reference's node statuses would give its P2 node fluid-solid DoFs
    that only one region holds."""
    r = np.linalg.norm(X, axis=1)
    on = np.abs(r - R) <= 1e-12 * R
    side = element_radius(X, T) < R
    cand = on[T].sum(axis=1) >= 2
    while True:
        ti = np.nonzero(cand)[0]
        edges = {}
        for t in ti:
            for p, q in _TET_EDGES:
                u, w = int(T[t, p]), int(T[t, q])
                if on[u] and on[w]:
                    key = (u, w) if u < w else (w, u)
                    edges.setdefault(key, []).append(t)
        chords = [(k, ts) for k, ts in edges.items() if len(set(side[ts].tolist())) == 1]
        if not chords:
            return X, T
        # a chord's ring of tetrahedra is complete when every tetrahedron on it
        # has two on-sphere ends; bisect the first chord, then look again
        (u, w), _ = chords[0]
        ring = np.nonzero(((T == u).any(axis=1)) & ((T == w).any(axis=1)))[0]
        m = X.shape[0]
        X = np.concatenate([X, 0.5 * (X[u] + X[w])[None, :]], axis=0)
        on = np.concatenate([on, [False]])
        new = []
        for t in ring:
            t1 = T[t].copy()
            t1[t1 == w] = m
            t2 = T[t].copy()
            t2[t2 == u] = m
            new += [t1, t2]
        keep = np.ones(T.shape[0], dtype=bool)
        keep[ring] = False
        sd = side[ring]
        T = np.concatenate([T[keep], np.asarray(new)], axis=0)
        side = np.concatenate([side[keep], np.repeat(sd, 2)])
        cand = np.concatenate([cand[keep], np.ones(len(new), dtype=bool)])
        T = _orient(X, T)


def _check_conforming(X, T, radii):
    """Every tetrahedron on one side of every interface sphere (vertices on
    the sphere count for both sides)."""
    rX = np.linalg.norm(X, axis=1)
    rc = element_radius(X, T)
    for R in radii:
        tol = 1e-9 * R
        inside = rc < R
        bad = (inside & (rX[T] > R + tol).any(axis=1)) | (~inside & (rX[T] < R - tol).any(axis=1))
        if bad.any():
            raise ValueError(f"{int(bad.sum())} tetrahedra cross the interface r = {R}: refine or change the seed")


def _promote(X, T, on_surf, porder):
    """P1 or P2 node set of the tetrahedra T over vertices X: P2 adds one node
    per unique edge at its midpoint (numbered after the vertices, in order of
    the sorted vertex pairs), in the reference's local slots."""
    if porder == 1:
        return X, T.copy(), on_surf
    if porder != 2:
        raise ValueError("pOrder must be 1 or 2")
    nv = X.shape[0]
    pairs = np.concatenate([np.sort(T[:, [a, b]], axis=1) for a, b, _ in _P2_EDGES], axis=0)
    key = pairs[:, 0] * nv + pairs[:, 1]
    ukey, inv = np.unique(key, return_inverse=True)
    ea, eb = ukey // nv, ukey % nv
    mid = 0.5 * (X[ea] + X[eb])
    nodes = np.concatenate([X, mid], axis=0)
    surf = np.concatenate([on_surf, on_surf[ea] & on_surf[eb]])
    elems = np.empty((T.shape[0], 10), dtype=np.int64)
    for slot, v in zip(_P2_VERT_SLOTS, range(4)):
        elems[:, slot] = T[:, v]
    ne = T.shape[0]
    for k, (_, _, slot) in enumerate(_P2_EDGES):
        elems[:, slot] = nv + inv[k * ne:(k + 1) * ne]
    return nodes, elems, surf


def _assemble_mesh(X, T, on_surf, porder, *, seed, h, reorder_elems):
    nodes, elems, surf = _promote(X, np.asarray(T, dtype=np.int64), on_surf, porder)
    order = _morton_order(nodes)
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    nodes = nodes[order]
    surf = surf[order]
    elems = rank[elems]
    tets = elems[:, list(_P2_VERT_SLOTS if porder == 2 else (0, 1, 2, 3))]
    if reorder_elems:
        # element order follows the Morton code of the first vertex
        eorder = np.lexsort((elems[:, -1], tets.min(axis=1)))
        elems = elems[eorder]
        tets = tets[eorder]
    return TetMesh(nodes=np.ascontiguousarray(nodes), elems=np.ascontiguousarray(elems.astype(np.int32)),
                   tets=np.ascontiguousarray(tets.astype(np.int32)), porder=porder, surface=surf,
                   seed=seed, h=h, vlist=order.astype(np.int64))


def mesh_from_vertices(X, T, porder: int, *, surface=None) -> TetMesh:
    """A TetMesh over given vertices X (nv, 3) and tetrahedra T (ne, 4),
    0-based, positively oriented.  Element order is kept (per-element model
    data stays aligned); nodes are renumbered in Morton order and ``vlist``
    records each node's id in the input numbering (vertices 0..nv-1, edge
    nodes nv.. in the order of ``_promote``)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.int64)
    vol = _tet_volumes(X, T)
    if np.any(vol <= 0.0):
        raise ValueError(f"{int(np.sum(vol <= 0))} tetrahedra are not positively oriented")
    on_surf = np.zeros(X.shape[0], dtype=bool) if surface is None else np.asarray(surface, dtype=bool)
    lrms = np.sqrt(np.mean([np.sum((X[T[:, b]] - X[T[:, a]]) ** 2, axis=1) for a, b, _ in _P2_EDGES], axis=0))
    return _assemble_mesh(X, T, on_surf, porder, seed=-1, h=float(np.mean(lrms)), reorder_elems=False)


def ellipsoid_mesh(n_tets: int, flattening: float = 1.0 / 300.0, **kw) -> TetMesh:
    """Oblate ellipsoid: the ball mesh scaled by (1, 1, 1 - f) (volume-preserving scale not applied)."""
    return ball_mesh(n_tets, axes=(1.0, 1.0, 1.0 - flattening), **kw)
