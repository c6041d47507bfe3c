"""CPU pins of the fluid-solid producer (CGFSE3D_ISO on the layout of
matrixstruct_general, src/mod_cg_create_matrix.f90:103-977, 1458-2033).

What pins it, without the reference's own build:
- exact reference-face integrals (the P2 and P1 triangle mass matrices);
- structure: E = ET^T exactly, Ad and B exactly symmetric, B's diagonal
  positive, Ap negative definite, every fluid-solid interface face found;
- rigid-body motions without gravity: a translation or rotation of the whole
  body (both displacement triples of the fluid-solid nodes moved alike) is
  in the null space of Ad - E Ap^-1 ET to rounding.  ET u = 0 needs the
  interface term with the right sign and normal: planting a sign flip there
  breaks it;
- mass: u^T B u of a unit translation is sum(rho_bar * volume) exactly;
- gravity (JOB=2): a rigid rotation of a spherical hydrostatic body costs
  no energy, so its Rayleigh quotient tends to zero under refinement while a
  translation's stays finite;
- the mu -> 0 limit: the displacement formulation with a nearly liquid core
  reproduces the first multiplet to discretization accuracy.
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse.linalg as spl

R_EARTH = 6371.0


def _body(model_fn, n_tets, porder, seed=1019):
    from normalmodes_amd.mesh import ball_mesh
    m = model_fn()
    return m, ball_mesh(n_tets, porder=porder, interfaces=m.interfaces, scale=m.radius, seed=seed)


def _rigid(Q, f):
    """Nodal field f(x) on every displacement triple (both triples at the
    fluid-solid nodes)."""
    u = np.zeros(Q.n)
    X = Q.mesh.nodes
    for v in range(Q.mesh.n_nodes):
        fv = f(X[v])
        u[Q.vstt[v]:Q.vstt[v] + 3] = fv
        if Q.vstat[v] in (2, 5):
            u[Q.vstt[v] + 3:Q.vstt[v] + 6] = fv
    return u


def _schur(Q):
    Ad, E, ET = Q.Ad.to_scipy(), Q.E.to_scipy(), Q.ET.to_scipy()
    lu = spl.splu(Q.Ap.to_scipy().tocsc())
    return lambda u: Ad @ u - E @ lu.solve(ET @ u)


def test_face_mass_closed_form():
    """P2 triangle mass matrix (area A): A/180 [6 on vertices, 32 on edge
    nodes, -1 vertex-vertex, -4 vertex-opposite edge, 16 edge-edge, 0
    vertex-adjacent edge]; P1: A/12 [2 diagonal, 1 off]; reference face
    area 2 for every face, nodes opposite vertex f excluded."""
    from normalmodes_amd.refelem import reference_faces, vertex_slots
    fm, M = reference_faces(2)
    vs = vertex_slots(2)
    for f in range(4):
        assert vs[f] not in fm[f] and all(v in fm[f] for k, v in enumerate(vs) if k != f)
        verts = [i for i, node in enumerate(fm[f]) if node in vs]
        edges = [i for i in range(6) if i not in verts]
        assert abs(M[f].sum() - 2.0) < 1e-14
        for i in verts:
            assert abs(M[f][i, i] - 6 * 2 / 180) < 1e-15
            for j in verts:
                if j != i:
                    assert abs(M[f][i, j] + 2 / 180) < 1e-15
        for i in edges:
            assert abs(M[f][i, i] - 32 * 2 / 180) < 1e-15
            for j in edges:
                if j != i:
                    assert abs(M[f][i, j] - 16 * 2 / 180) < 1e-15
        # each vertex has one opposite edge (-4) and two adjacent ones (0)
        for i in verts:
            row = np.sort(M[f][i, edges] * 90)
            assert np.allclose(row, [-4, 0, 0], atol=1e-12)
    fm1, M1 = reference_faces(1)
    for f in range(4):
        assert np.allclose(M1[f] * 6, np.ones((3, 3)) + np.eye(3), atol=1e-15)


@pytest.fixture(scope="module", params=[(1, "fluid_core", 1400), (2, "fluid_core", 1400), (1, "three_layer", 12000)],
                ids=["p1-core", "p2-core", "p1-3layer"])
def body(request):
    from normalmodes_amd import models
    from normalmodes_amd.assemble import assemble
    porder, name, nt = request.param
    fn = (lambda: models.fluid_core(0.5, radius=R_EARTH)) if name == "fluid_core" else \
        (lambda: models.three_layer(R_EARTH))
    m, me = _body(fn, nt, porder)
    return m, me, assemble(me, m, job=1)


def test_structure(body):
    from normalmodes_amd.assemble import FluidProblem
    m, me, Q = body
    assert isinstance(Q, FluidProblem)
    E, ET = Q.E.to_scipy(), Q.ET.to_scipy()
    assert (E - ET.T).nnz == 0 or abs(E - ET.T).max() == 0.0
    Ad, B = Q.Ad.to_scipy(), Q.B.to_scipy()
    assert abs(Ad - Ad.T).max() == 0.0 and abs(B - B.T).max() == 0.0
    assert B.diagonal().min() > 0
    # six DoFs at fluid-solid nodes, one pressure per node of status 1, 2, 4, 5
    six = np.isin(Q.vstat, (2, 5))
    assert Q.n == 3 * me.n_nodes + 3 * int(six.sum())
    assert Q.n_p == int(np.isin(Q.vstat, (1, 2, 4, 5)).sum())
    # Ap negative definite (Cholesky of -Ap succeeds; its largest eigenvalue < 0)
    Ap = Q.Ap.to_scipy().toarray()
    np.linalg.cholesky(-Ap)
    assert Q.stats["interface_elements"] > 0 and Q.stats["four_fs_vertices"] == 0
    # every fluid-solid vertex lies on an interface sphere
    rv = np.linalg.norm(me.nodes[Q.vstat == 2], axis=1) / m.radius
    assert np.min(np.abs(rv[:, None] - np.array(m.interfaces)[None, :]), axis=1).max() < 1e-12


def test_rigid_motions_null_space(body):
    m, me, Q = body
    S = _schur(Q)
    ET = Q.ET.to_scipy()
    scale = np.abs(Q.Ad.val).max()
    for f in (lambda x: np.array([1.0, 0, 0]), lambda x: np.array([0, 0, 1.0]),
              lambda x: np.cross([0, 0, 1.0], x), lambda x: np.cross([0.6, 0.8, 0], x)):
        u = _rigid(Q, f)
        assert np.linalg.norm(S(u)) <= 1e-14 * scale * np.linalg.norm(u)
        assert np.linalg.norm(ET @ u) <= 1e-14 * np.abs(ET).max() * np.linalg.norm(u)
    # a planted mistake: the interface coupling (ET entries in solid-triple
    # columns) with the wrong sign no longer annihilates a translation
    bad = ET.copy().tocsr()
    solid_cols = np.zeros(Q.n, dtype=bool)
    for v in np.nonzero(np.isin(Q.vstat, (2, 5)))[0]:
        solid_cols[Q.vstt[v]:Q.vstt[v] + 3] = True
    bad.data[solid_cols[bad.indices]] *= -1.0
    u = _rigid(Q, lambda x: np.array([1.0, 0, 0]))
    assert np.linalg.norm(bad @ u) > 1e-3 * np.abs(ET).max() * np.linalg.norm(u)


def test_total_mass(body):
    from normalmodes_amd.mesh import _tet_volumes
    m, me, Q = body
    u = _rigid(Q, lambda x: np.array([0, 1.0, 0]))
    mass = u @ (Q.B.to_scipy() @ u)
    vol = np.abs(_tet_volumes(me.nodes, me.tets.astype(np.int64)))
    exact = float(np.sum(Q.rho.mean(axis=1) * vol))
    assert abs(mass - exact) <= 1e-13 * exact


def test_gravity_rotation_tends_to_zero_mode():
    """JOB=2: Rayleigh quotient of a rigid rotation -> 0 under refinement,
    far below a translation's."""
    from normalmodes_amd import models
    from normalmodes_amd.assemble import assemble
    out = []
    for nt in (1400, 5000):
        m, me = _body(lambda: models.fluid_core(0.5, radius=R_EARTH), nt, 2)
        Q = assemble(me, m, job=2)
        assert Q.g0 is not None
        S = _schur(Q)
        B = Q.B.to_scipy()

        def rq(f):
            u = _rigid(Q, f)
            return (u @ S(u)) / (u @ (B @ u))
        out.append((abs(rq(lambda x: np.cross([0, 0, 1.0], x))), rq(lambda x: np.array([1.0, 0, 0]))))
    (r1, t1), (r2, t2) = out
    assert t1 > 0 and t2 > 0 and abs(t1 - t2) < 0.05 * t1
    assert r1 < 0.05 * t1 and r2 < 0.6 * r1


def test_mu_to_zero_limit():
    """A nearly liquid core (vs = 1e-3) in the displacement formulation gives
    the fluid-solid pencil's first multiplet to discretization accuracy
    (constant density, so the fluid-solid path's element-average mass equals
    the solid path's nodal one)."""
    from normalmodes_amd.assemble import assemble
    from normalmodes_amd.mesh import ball_mesh
    from normalmodes_amd.models import Layer, RadialModel

    def mdl(vcore):
        return RadialModel([Layer(0.0, 0.5, (5.5, 5.5), (10.0, 10.0), (vcore, vcore)),
                            Layer(0.5, 1.0, (5.5, 5.5), (10.0, 10.0), (6.0, 6.0))], radius=R_EARTH)
    me = ball_mesh(900, porder=2, interfaces=(0.5,), scale=R_EARTH)
    Q = assemble(me, mdl(0.0), job=1)
    Ad, E, ET = (Q.Ad.to_scipy().toarray(), Q.E.to_scipy().toarray(), Q.ET.to_scipy().toarray())
    S = Ad - E @ np.linalg.solve(Q.Ap.to_scipy().toarray(), ET)
    w = sl.eigh(0.5 * (S + S.T), Q.B.to_scipy().toarray(), eigvals_only=True)
    f = np.sqrt(np.abs(w)) / (2 * np.pi) * 1e3
    first = f[f > 0.05][:5]
    P = assemble(me, mdl(1e-3), job=1)
    w2 = sl.eigh(P.A.to_scipy().toarray(), P.B.to_scipy().toarray(), eigvals_only=True)
    f2 = np.sqrt(np.abs(w2)) / (2 * np.pi) * 1e3
    near = np.array([f2[np.argmin(np.abs(f2 - x))] for x in first])
    assert np.max(np.abs(near - first) / first) < 2e-3
    # the rigid modes and the fluid's zero-energy displacements sit at zero
    assert np.sum(f < 1e-3) > 6


def test_node_status_rules():
    """Edge nodes take the smaller status of their vertices; an edge between
    two interface vertices that only one region holds would leave empty rows
    in the reference's matrices, and is rejected."""
    from normalmodes_amd import models
    from normalmodes_amd.assemble import node_status, face_neighbours
    from normalmodes_amd.models import element_node_values
    m, me = _body(lambda: models.fluid_core(0.5, radius=R_EARTH), 1400, 2)
    rho, vp, vs = element_node_values(me, m)
    fluid = vs.max(axis=1) < 1e-6
    st = node_status(me, fluid)
    assert set(np.unique(st[me.tets.ravel()])) <= {0, 1, 2}
    edge_nodes = np.setdiff1d(np.unique(me.elems), np.unique(me.tets))
    assert set(np.unique(st[edge_nodes])) <= {3, 4, 5}
    # relabel one solid element touching the interface as fluid: some of its
    # edges now join two interface vertices through one region only
    ne = face_neighbours(me)
    assert ne.shape == (me.n_elem, 4) and (ne == -1).sum() > 0
    cand = np.nonzero(~fluid & (st[me.tets] == 2).any(axis=1))[0]
    bad = fluid.copy()
    bad[cand[:40]] = True
    with pytest.raises(ValueError):
        node_status(me, bad)
