"""The CPU producer of A and B: reference element, element formulas and CSR
layout, checked against exact integrals, invariants, the committed fixtures
and a closed-form mode of the homogeneous elastic sphere."""
import json
from pathlib import Path

import numpy as np
import pytest

from normalmodes_amd.refelem import reference_element

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_reference_element_exact():
    _, M1, D1 = reference_element(1)
    # P1 mass on the reference tet (volume 4/3): (1 + delta_ij) / 15
    assert np.allclose(M1 * 15.0, np.ones((4, 4)) + np.eye(4), atol=1e-15)
    for p in (1, 2):
        nodes, M, D = reference_element(p)
        assert abs(M.sum() - 4.0 / 3.0) < 1e-14            # int 1 = volume
        assert np.allclose(M, M.T, atol=0)                  # exactly symmetric
        assert np.abs(D.sum(axis=2)).max() < 1e-13          # derivative of a constant
        for d in range(3):                                  # coordinate functions
            for e in range(3):
                assert np.abs(D[d] @ nodes[:, e] - (1.0 if d == e else 0.0)).max() < 1e-13
        # exact for quadratics: d/dr (r^2) = 2 r (P2)
        if p == 2:
            r = nodes[:, 0]
            assert np.abs(D[0] @ (r * r) - 2 * r).max() < 1e-13
        assert np.linalg.eigvalsh(M).min() > 0


def test_p2_node_order_matches_reference_layout():
    nodes, _, _ = reference_element(2)
    V = {(-1, -1, -1): 0, (1, -1, -1): 2, (-1, 1, -1): 5, (-1, -1, 1): 9}
    for xyz, slot in V.items():
        assert np.allclose(nodes[slot], xyz)
    # edge (1-2) -> slot 1 at the midpoint of vertices 1 and 2
    assert np.allclose(nodes[1], [0, -1, -1]) and np.allclose(nodes[8], [-1, 0, 0])


def _rigid_modes(X):
    R = []
    for c in range(3):
        r = np.zeros_like(X)
        r[:, c] = 1.0
        R.append(r.ravel())
    for c in range(3):
        a, b = (c + 1) % 3, (c + 2) % 3
        r = np.zeros_like(X)
        r[:, a] = -X[:, b]
        r[:, b] = X[:, a]
        R.append(r.ravel())
    return R


@pytest.mark.parametrize("fixture", ["small_p2", "small_p1", "layered_p2"])
def test_invariants(fixture, request):
    P = request.getfixturevalue(fixture)
    A, B = P.A.to_scipy(), P.B.to_scipy()
    assert abs(A - A.T).max() == 0.0 and abs(B - B.T).max() == 0.0
    scale = abs(A).max()
    for r in _rigid_modes(P.mesh.nodes):                    # affine fields are in the space
        assert np.abs(A @ r).max() < 1e-12 * scale * np.abs(r).max()
    # total mass of one component = int rho over the body
    e = np.zeros(P.n)
    e[0::3] = 1.0
    vol_rho = e @ (B @ e)
    X = P.mesh.nodes
    T = P.mesh.tets
    v = np.einsum("ij,ij->i", X[T[:, 1]] - X[T[:, 0]], np.cross(X[T[:, 2]] - X[T[:, 0]], X[T[:, 3]] - X[T[:, 0]])) / 6
    assert v.min() > 0
    if P.rho.min() == P.rho.max():
        assert abs(vol_rho - P.rho[0, 0] * v.sum()) < 1e-12 * vol_rho
    # B = M (x) I3: only same-component columns, identical values per component
    for i in range(0, 30, 3):
        row = [B.getrow(i + p) for p in range(3)]
        assert all(np.array_equal(row[0].data, row[p].data) for p in range(3))
        assert all(np.all(row[p].indices % 3 == p) for p in range(3))
    # positive definite mass, positive semi-definite stiffness (spot check)
    x = np.random.default_rng(0).standard_normal(P.n)
    assert x @ (B @ x) > 0 and x @ (A @ x) > 0


def test_tiny_fixture_checksums(small_p2, small_p1):
    g = json.loads((GOLDEN / "tiny.json").read_text())
    for key, P in (("p2", small_p2), ("p1", small_p1)):
        G = g[key]
        assert P.n == G["n_dof"] and P.A.nnz == G["nnz_A"] and P.B.nnz == G["nnz_B"]
        assert abs(np.abs(P.A.val).sum() - G["sum_abs_A"]) < 1e-10 * G["sum_abs_A"]
        assert abs(P.B.val.sum() - G["sum_B"]) < 1e-12 * G["sum_B"]
        assert abs(np.linalg.norm(P.A.val) - G["fro_A"]) < 1e-12 * G["fro_A"]


def test_c1_lowest_multiplet_near_lamb_toroidal():
    """The lowest non-rigid five-fold multiplet of the C1 ball is 0T2 of a
    homogeneous sphere: x j_2'(x) - j_2(x) = 0, i.e. j_2(x) = x j_3(x) with
    x = omega R / vs; lambda = x^2 for R = vs = 1 (2.5011^2 = 6.2557).  The P2
    discretisation error at ~12k tets is below 1%, and the multiplet is
    split by the mesh at the 1e-3 level only."""
    from scipy.optimize import brentq
    from scipy.special import spherical_jn as jn
    x = brentq(lambda t: jn(2, t) - t * jn(3, t), 1.5, 3.5)
    g = json.loads((GOLDEN / "c1_eigs.json").read_text())
    w = np.array(g["eigenvalues"])
    assert np.all(np.abs(w[:6]) < 1e-9)                    # rigid modes
    mult = w[6:11]
    assert np.ptp(mult) / mult.mean() < 1e-3
    assert 0 < mult.mean() / x ** 2 - 1 < 0.01              # FE bound from above


def test_fluid_body_needs_conforming_interfaces():
    """A layered body with a fluid layer assembles into the fluid-solid blocks
    when its mesh has the layer boundaries as tetrahedron faces; on a mesh
    cut across them the reference's node statuses leave empty rows, and the
    producer refuses it."""
    from normalmodes_amd.assemble import FluidProblem, assemble
    from normalmodes_amd.mesh import ball_mesh
    from normalmodes_amd.models import three_layer
    with pytest.raises(ValueError):
        assemble(ball_mesh(600, porder=2), three_layer())
    m = three_layer()
    Q = assemble(ball_mesh(12000, porder=1, interfaces=m.interfaces), m)
    assert isinstance(Q, FluidProblem) and Q.n_p > 0
