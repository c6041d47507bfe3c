"""The sliver-free layered-body mesh (body-centred cubic lattice cut by the
interface spheres with isosurface stuffing): dihedral-angle bounds, exact
conformity to every sphere, the ball's volume, consistent splits of the cut
pieces, and the fluid-solid producer on it."""
import numpy as np
import pytest

R_EARTH = 6371.0


def _dihedral(X, T):
    out = []
    for a, b in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)):
        o = [i for i in range(4) if i not in (a, b)]
        e = X[T[:, b]] - X[T[:, a]]
        e /= np.linalg.norm(e, axis=1)[:, None]
        u = X[T[:, o[0]]] - X[T[:, a]]
        w = X[T[:, o[1]]] - X[T[:, a]]
        u -= (u * e).sum(1)[:, None] * e
        w -= (w * e).sum(1)[:, None] * e
        c = (u * w).sum(1) / np.linalg.norm(u, axis=1) / np.linalg.norm(w, axis=1)
        out.append(np.degrees(np.arccos(np.clip(c, -1.0, 1.0))))
    return np.stack(out, 1)


@pytest.fixture(scope="module", params=[3000, 20000])
def layered(request):
    from normalmodes_amd import models
    from normalmodes_amd.mesh import stuffed_ball_mesh
    m = models.three_layer(R_EARTH)
    return m, stuffed_ball_mesh(request.param, porder=2, interfaces=m.interfaces, scale=m.radius)


def test_dihedral_bounds_and_volume(layered):
    from normalmodes_amd.mesh import _tet_volumes
    m, me = layered
    X, T = me.nodes, me.tets.astype(np.int64)
    D = _dihedral(X, T)
    # the lattice's own tetrahedra have 60 and 90 degree angles; the cut
    # pieces stay within [18, 150] degrees (isosurface stuffing's bound on
    # the BCC lattice is [10.7, 164.8])
    assert D.min() > 15.0 and D.max() < 155.0, (D.min(), D.max())
    share = np.mean(np.all(np.isclose(D, 60.0, atol=1e-6) | np.isclose(D, 90.0, atol=1e-6), axis=1))
    # most tetrahedra away from the three spheres are untouched lattice ones
    # (about 12% at 3k tetrahedra, 48% at 20k, 68% at 100k)
    assert share > (0.1 if me.n_elem < 10000 else 0.4), share
    vol = _tet_volumes(X, T)
    assert vol.min() > 0
    ball = 4.0 / 3.0 * np.pi * m.radius ** 3
    # inscribed polyhedron: below the ball's volume by the chord error O(h^2)
    h = me.h
    assert 0 < 1 - vol.sum() / ball < 3 * (h / m.radius) ** 2
    rv = np.linalg.norm(X[np.unique(T)], axis=1)
    assert rv.max() <= m.radius * (1 + 1e-12)
    vs = np.unique(T)
    sv = vs[me.surface[vs]]                      # surface vertices (edge nodes sit on chords)
    assert sv.size > 0
    assert np.all(np.abs(np.linalg.norm(X[sv], axis=1) - m.radius) <= 1e-9 * m.radius)


def test_interfaces_are_tetrahedron_faces(layered):
    """Every tetrahedron lies on one side of every interface (vertices on the
    sphere count for both), and each interface sphere carries a closed surface
    of faces shared by one tetrahedron on each side."""
    from normalmodes_amd.mesh import _check_conforming, element_radius
    m, me = layered
    X, T = me.nodes, me.tets.astype(np.int64)
    radii = [f * m.radius for f in m.interfaces]
    _check_conforming(X, T, radii)
    rc = element_radius(X, T)
    for R in radii:
        on = np.abs(np.linalg.norm(X, axis=1) - R) <= 1e-9 * R
        faces = {}
        for f in ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)):
            F = np.sort(T[:, f], axis=1)
            sel = on[F].all(axis=1)
            for t, key in zip(np.nonzero(sel)[0], map(tuple, F[sel])):
                faces.setdefault(key, []).append(t)
        iface = [ts for ts in faces.values() if len(ts) == 2 and (rc[ts[0]] < R) != (rc[ts[1]] < R)]
        assert iface
        # closed: every edge of the interface surface is shared by two of its faces
        keys = [k for k, ts in faces.items() if len(ts) == 2 and (rc[ts[0]] < R) != (rc[ts[1]] < R)]
        from collections import Counter
        cnt = Counter()
        for a, b, c in keys:
            cnt[(a, b)] += 1
            cnt[(a, c)] += 1
            cnt[(b, c)] += 1
        assert set(cnt.values()) == {2}
        # total area close to the sphere's (inscribed: slightly below)
        A = sum(0.5 * np.linalg.norm(np.cross(X[b] - X[a], X[c] - X[a])) for a, b, c in keys)
        assert 0 < 1 - A / (4 * np.pi * R * R) < 0.4 * (me.h / R) ** 2


def test_prism_and_pyramid_splits_consistent():
    """The three tetrahedra of a prism and the two of a pyramid fill the
    solid exactly, and each quadrilateral face is split through its smallest
    vertex id whichever solid it belongs to."""
    from normalmodes_amd.mesh import _prism, _pyramid, _tet_volumes, _orient
    rng = np.random.default_rng(3)
    for trial in range(50):
        ids = rng.permutation(60)[:6].tolist()
        X = np.zeros((60, 3))
        base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float) + rng.uniform(-0.1, 0.1, (3, 3))
        top = base + np.array([0.1, -0.05, 1.0]) + rng.uniform(-0.1, 0.1, (3, 3))
        X[ids[:3]] = base
        X[ids[3:]] = top
        T = _orient(X, np.array(_prism(ids[:3], ids[3:])))
        v = _tet_volumes(X, T)
        assert np.all(v > 0)
        # volume of the (possibly non-planar-faced) prism = sum of the split
        # with the same face diagonals; compare with an independent split
        # through the centroid of the six vertices
        for q in ((0, 1), (1, 2), (2, 0)):
            quad = [ids[q[0]], ids[q[1]], ids[3 + q[1]], ids[3 + q[0]]]
            lo = min(quad)
            # the tetrahedra's faces on this quad use the diagonal through lo
            tri = [set(f) for t in T for f in ({t[0], t[1], t[2]}, {t[0], t[1], t[3]}, {t[0], t[2], t[3]},
                                                {t[1], t[2], t[3]}) if f <= set(quad)]
            assert len(tri) == 2 and all(lo in f for f in tri)
        apex = ids[5]
        quad = ids[:4]
        Xp = np.zeros((60, 3))
        Xp[quad] = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
        Xp[apex] = [0.4, 0.6, 0.8]
        Tp = _orient(Xp, np.array(_pyramid(apex, quad)))
        assert abs(_tet_volumes(Xp, Tp).sum() - 0.8 / 3) < 1e-12
        tri = [set(t[1:]) if t[0] == apex else None for t in _pyramid(apex, quad)]
        assert all(min(quad) in f for f in tri)


def test_warping_prevents_slivers(monkeypatch):
    """A planted mistake: with the warp thresholds at zero, cut points near
    lattice vertices leave thin pieces (the bounds above fail)."""
    from normalmodes_amd import mesh as M
    from normalmodes_amd import models
    m = models.three_layer(R_EARTH)
    monkeypatch.setattr(M, "STUFF_ALPHA_LONG", 0.0)
    monkeypatch.setattr(M, "STUFF_ALPHA_SHORT", 0.0)
    try:
        me = M.stuffed_ball_mesh(20000, porder=1, interfaces=m.interfaces, scale=m.radius)
    except ValueError:
        return                                   # degenerate piece: refused
    D = _dihedral(me.nodes, me.tets.astype(np.int64))
    assert D.min() < 10.0 or D.max() > 160.0


def test_fluid_solid_producer_on_stuffed_mesh():
    """The fluid-solid pencil on the stuffed body: node statuses resolve (no
    chord edge through one region), rigid motions in the null space of the
    Schur complement without gravity, total mass exact."""
    import scipy.sparse.linalg as spl
    from normalmodes_amd import models
    from normalmodes_amd.assemble import FluidProblem, assemble
    from normalmodes_amd.mesh import stuffed_ball_mesh, _tet_volumes
    m = models.three_layer(R_EARTH)
    me = stuffed_ball_mesh(6000, porder=2, interfaces=m.interfaces, scale=m.radius)
    Q = assemble(me, m, job=1)
    assert isinstance(Q, FluidProblem) and Q.n_p > 0 and Q.stats["four_fs_vertices"] == 0
    Ad, E, ET = Q.Ad.to_scipy(), Q.E.to_scipy(), Q.ET.to_scipy()
    lu = spl.splu(Q.Ap.to_scipy().tocsc())
    X = me.nodes
    scale = np.abs(Q.Ad.val).max()
    for f in (lambda x: np.array([0, 1.0, 0]), lambda x: np.cross([0.3, 0, 1.0], x)):
        u = np.zeros(Q.n)
        for v in range(me.n_nodes):
            fv = f(X[v])
            u[Q.vstt[v]:Q.vstt[v] + 3] = fv
            if Q.vstat[v] in (2, 5):
                u[Q.vstt[v] + 3:Q.vstt[v] + 6] = fv
        S = Ad @ u - E @ lu.solve(ET @ u)
        assert np.linalg.norm(S) <= 1e-14 * scale * np.linalg.norm(u)
    u = np.zeros(Q.n)
    for v in range(me.n_nodes):
        u[Q.vstt[v]] = 1.0
        if Q.vstat[v] in (2, 5):
            u[Q.vstt[v] + 3] = 1.0
    vol = _tet_volumes(X, me.tets.astype(np.int64))
    exact = float(np.sum(Q.rho.mean(axis=1) * vol))
    assert abs(u @ (Q.B.to_scipy() @ u) - exact) <= 1e-13 * exact
