"""Fluid-solid Schur operator (sparsefsAV) on the device.

The finite-element producer assembles solids only, so the pencil is synthetic
but has the reference's block structure (src/mod_matvec.f90:98-222): Ad
(solid stiffness), Ap = -(pressure mass) negative definite, ET coupling
pressure to displacements on the fluid nodes' stencils, E = ET^T.  The
oracle is the dense Schur complement D (Ad - E Ap^-1 ET) D."""
import numpy as np
import pytest
import scipy.sparse as sp


def synthetic_fluid_problem(P, seed=3, r_fluid=0.55, coupling=10.0):
    from normalmodes_amd.assemble import CSR
    from normalmodes_amd.driver import FluidProblem
    rng = np.random.default_rng(seed)
    B = P.B.to_scipy().tocsr()
    M = B[0::3, 0::3].tocsr()                          # scalar mass (B = M (x) I3)
    r = np.linalg.norm(P.mesh.nodes, axis=1)
    f = np.where(r < r_fluid)[0]
    Ap = (-2.0 * M[f][:, f]).tocsr()
    g = rng.standard_normal((P.mesh.n_nodes, 3)) * coupling
    Mf = M[f].tocoo()                                   # pressure row p couples to stencil nodes w
    rows = np.repeat(Mf.row, 3)
    cols = (3 * Mf.col[:, None] + np.arange(3)[None, :]).ravel()
    vals = (Mf.data[:, None] * g[Mf.col]).ravel()
    ET = sp.csr_matrix((vals, (rows, cols)), shape=(f.size, P.n))
    ET.sort_indices()
    E = ET.T.tocsr()
    E.sort_indices()

    def csr(S):
        S = S.tocsr()
        S.sort_indices()
        return CSR(S.shape[0], S.shape[1], S.indptr.astype(np.int64), S.indices.astype(np.int32),
                   S.data.astype(np.float64))
    return FluidProblem(Ad=P.A, E=csr(E), ET=csr(ET), Ap=csr(Ap), B=P.B, porder=P.mesh.porder, name="synthetic-fs")


def dense_schur(Q):
    Ad = Q.Ad.to_scipy().toarray()
    E = Q.E.to_scipy().toarray()
    ET = Q.ET.to_scipy().toarray()
    Ap = Q.Ap.to_scipy().toarray()
    return Ad - E @ np.linalg.solve(Ap, ET)


def test_synthetic_structure(small_p2):
    Q = synthetic_fluid_problem(small_p2)
    assert 0 < Q.n_p < small_p2.n // 3
    Ap = Q.Ap.to_scipy().toarray()
    assert np.all(np.linalg.eigvalsh(Ap) < 0)
    S = dense_schur(Q)
    assert np.allclose(S, S.T, atol=1e-10 * np.abs(S).max())
    # the coupling term is positive semi-definite and not negligible
    w = np.linalg.eigvalsh(S - Q.Ad.to_scipy().toarray())
    assert w.min() > -1e-9 * w.max() and w.max() > 1e-2 * np.abs(Q.Ad.val).max()


@pytest.fixture(scope="module")
def fs_setup(small_p2):
    from normalmodes_amd.driver import setupmatvec_fs
    Q = synthetic_fluid_problem(small_p2)
    mv = setupmatvec_fs(Q, bdeg="auto")
    yield Q, mv
    mv.free()


@pytest.mark.gpu
def test_fs_operator_matches_dense_schur(fs_setup):
    from normalmodes_amd import pevsl as pv
    from normalmodes_amd.driver import _open_solver
    Q, mv = fs_setup
    S = dense_schur(Q)
    St = S * mv.d[:, None] * mv.d[None, :]
    # Dp and Ap~ = Dp (-Ap) Dp: Apdiagscaling, unit diagonal
    Ap = Q.Ap.to_scipy().toarray()
    assert np.allclose(mv.fs.dp, 1.0 / np.sqrt(-np.diag(Ap)), rtol=1e-14)
    assert mv.fs.apdeg == 100
    rng = np.random.default_rng(1)
    x = rng.standard_normal(mv.n)
    X = rng.standard_normal((4, mv.n))
    h = _open_solver(mv, precision=pv.FP64, seed=7, rtol=1e-9)
    try:
        y = pv.apply(h, 0, x)
        Y = pv.apply_block(h, 0, X)
    finally:
        pv.pevsl_finish(h)
    ref = St @ x
    assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)
    refY = X @ St.T
    assert np.linalg.norm(Y - refY) <= 1e-12 * np.linalg.norm(refY)


@pytest.mark.gpu
@pytest.mark.parametrize("block", [1, 4])
def test_fs_eigenpairs_match_dense(fs_setup, block):
    import scipy.linalg as sl
    from normalmodes_amd.driver import apply_pevsl
    Q, mv = fs_setup
    S = dense_schur(Q)
    B = Q.B.to_scipy().toarray()
    w = sl.eigh(S, B, eigvals_only=True)
    nz = w[w > 1e-6 * w.max()]
    a, b = 0.5 * (nz[2] + nz[3]), 0.5 * (nz[17] + nz[18])
    ref = w[(w >= a) & (w <= b)]
    r = apply_pevsl(mv, interval=(a, b), block=block)
    assert r.lam.size == ref.size
    assert np.max(np.abs(r.lam - ref) / ref) < 1e-10
    assert np.all(r.res2 <= 1e-10)


@pytest.mark.gpu
def test_fs_slices(fs_setup):
    import scipy.linalg as sl
    from normalmodes_amd.driver import apply_slices
    Q, mv = fs_setup
    w = sl.eigh(dense_schur(Q), Q.B.to_scipy().toarray(), eigvals_only=True)
    nz = w[w > 1e-6 * w.max()]
    a, b = 0.5 * (nz[2] + nz[3]), 0.5 * (nz[17] + nz[18])
    ref = w[(w >= a) & (w <= b)]
    r = apply_slices(mv, interval=(a, b), n_slices=2, slicer="uniform")
    assert r.lam.size == ref.size and np.max(np.abs(r.lam - ref) / ref) < 1e-10
