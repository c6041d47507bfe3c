"""Seeded tetrahedral meshes of a ball and their P1/P2 node sets.

The reference reads TetGen meshes produced by an external builder
(PlanetaryModels; README.md:41) through ``pnm_read_ele``/``pnm_read_node``
(src/mod_geometry.f90:140-425) and adds P2 edge nodes in ``pnm_add_edges``
(src/mod_geometry.f90:428-689).  This module replaces that producer with a
deterministic generator: a cubic lattice of spacing h with uniform jitter
+-jitter*h, points on the bounding sphere, Delaunay tetrahedralisation (Qhull),
removal of flat hull slivers, positive orientation, P2 edge nodes at edge
midpoints and a Morton (Z-order) node numbering for memory locality of the
assembled CSR rows.

Element-node layout follows the reference (see ``refelem``): for P2 the local
slots are vertices (0,2,5,9) and edges (1-2)->1, (1-3)->3, (2-3)->4, (1-4)->6,
(2-4)->7, (3-4)->8 (0-based slots, 1-based vertex names).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["TetMesh", "ball_mesh", "ellipsoid_mesh", "h_for_tets", "mesh_from_vertices"]

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


def _check_conforming(X, T, radii):
    """Every tetrahedron on one side of every interface sphere (vertices on
    the sphere count for both sides)."""
    rX = np.linalg.norm(X, axis=1)
    rc = np.linalg.norm(X[T].mean(axis=1), axis=1)
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
