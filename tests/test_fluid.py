"""Fluid-solid Schur operator (sparsefsAV, src/mod_matvec.f90:98-222, 498-520)
on the device.

Two kinds of pencils: synthetic ones with the reference's block structure
(Ad, Ap = -(pressure mass) negative definite, ET coupling pressure to the
fluid nodes' stencils, E = ET^T), and the assembled fluid-solid bodies of the
producer (CGFSE3D_ISO).  The reference for both is the dense *same operator*
the device applies: D (Ad + E Dp q(Ap~) Dp ET) D, where q is the Chebyshev
polynomial of the Ap~ solve (degree 100 at P2, 25 at P1) on the device's own
Lanczos bounds of Ap~ = Dp (-Ap) Dp, in the closed form pinned in
tests/test_oracle_pins.py (1 - t q(t) = T_{c+1}((th - t)/de) / T_{c+1}(th/de))."""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

from oracle import oracle as orc


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


def cheb_poly_values(t, lmin, lmax, deg):
    """q(t) of the Chebyshev semi-iteration after ``deg`` products on
    [lmin, lmax]: 1 - t q(t) = T_{deg+1}((th - t)/de) / T_{deg+1}(th/de)."""
    th, de = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    k = deg + 1
    T = np.polynomial.chebyshev.chebval
    R = T((th - t) / de, np.eye(k + 1)[k]) / T(th / de, np.eye(k + 1)[k])
    return (1.0 - R) / t


def same_operator(Q, mv):
    """The unscaled operator the device applies, Ad + E Dp q(Ap~) Dp ET,
    symmetrised, with q on the device's bounds and degree of the Ap~ solve."""
    Ad = Q.Ad.to_scipy().toarray()
    E = Q.E.to_scipy().toarray()
    ET = Q.ET.to_scipy().toarray()
    Ap = Q.Ap.to_scipy().toarray()
    dp = 1.0 / np.sqrt(-np.diag(Ap))
    Apt = -Ap * dp[:, None] * dp[None, :]
    th, U = np.linalg.eigh(0.5 * (Apt + Apt.T))
    q = cheb_poly_values(th, mv.fs.aplmin, mv.fs.aplmax, mv.fs.apdeg)
    C = (U * q[None, :]) @ U.T
    S = Ad + (E * dp[None, :]) @ C @ (dp[:, None] * ET)
    return 0.5 * (S + S.T)


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
    r = apply_pevsl(mv, interval=(a, b), block=block)
    # reference: shift-invert on the same operator (dense LAPACK leaves
    # absolute errors ~eps lam_max, above 1e-12 relative at the low end)
    S2 = same_operator(Q, mv)
    ref = orc.shift_invert_eigs(S2, B, a, b)
    assert r.lam.size == ref.size == w[(w >= a) & (w <= b)].size
    err = np.max(np.abs(r.lam - ref) / ref)
    print(f"synthetic fs block {block}: {ref.size} pairs, max rel err {err:.2e}", flush=True)
    assert err < 1e-12
    assert np.all(r.res2 <= 1e-10)
    # q(Ap~) = Ap~^-1 to rounding at degree 100 on this pencil: the exact
    # Schur complement has the same eigenvalues
    ref_exact = orc.shift_invert_eigs(S, B, a, b)
    assert np.max(np.abs(ref_exact - ref) / ref) < 1e-13


@pytest.mark.gpu
def test_fs_slices(fs_setup):
    import scipy.linalg as sl
    from normalmodes_amd.driver import apply_slices
    Q, mv = fs_setup
    B = Q.B.to_scipy().toarray()
    w = sl.eigh(dense_schur(Q), B, eigvals_only=True)
    nz = w[w > 1e-6 * w.max()]
    a, b = 0.5 * (nz[2] + nz[3]), 0.5 * (nz[17] + nz[18])
    ref = orc.shift_invert_eigs(same_operator(Q, mv), B, a, b)
    r = apply_slices(mv, interval=(a, b), n_slices=2, slicer="uniform")
    assert r.lam.size == ref.size and np.max(np.abs(r.lam - ref) / ref) < 1e-12


# ---------------------------------------------------------------- assembled bodies
ASSEMBLED = {
    # name: (model, n_tets, pOrder, job)
    "core-p2-job1": ("fluid_core", 900, 2, 1),
    "core-p2-job2": ("fluid_core", 900, 2, 2),
    "3layer-p1-job2": ("three_layer", 9000, 1, 2),
}


@pytest.fixture(scope="module", params=list(ASSEMBLED), ids=list(ASSEMBLED))
def assembled(request):
    from normalmodes_amd import models
    from normalmodes_amd.assemble import assemble
    from normalmodes_amd.driver import setupmatvec_fs
    from normalmodes_amd.mesh import ball_mesh
    name, nt, porder, job = ASSEMBLED[request.param]
    m = models.fluid_core(0.5, radius=6371.0) if name == "fluid_core" else models.three_layer(6371.0)
    me = ball_mesh(nt, porder=porder, interfaces=m.interfaces, scale=m.radius)
    Q = assemble(me, m, job=job)
    print(f"assembled {request.param}: {Q.n} DoF, {Q.n_p} pressures", flush=True)
    mv = setupmatvec_fs(Q, bdeg="auto")
    S = same_operator(Q, mv)
    B = Q.B.to_scipy().toarray()
    w = sl.eigh(S, B, eigvals_only=True)
    yield request.param, Q, mv, S, B, w
    mv.free()


def _seismic_interval(w, n_modes=20):
    """An interval over the first ~n_modes eigenvalues above the zero-energy
    and core-mode cluster, with its ends in gaps."""
    f = np.sign(w) * np.sqrt(np.abs(w))
    top = f.max()
    i0 = int(np.argmax(f > 4e-3 * top))                # first mode above the cluster
    i1 = min(i0 + n_modes, w.size - 2)
    g = np.diff(w)
    j0 = i0 + int(np.argmax(g[i0:i0 + 6]))              # widest gap among the first few
    j1 = i1 + int(np.argmax(g[i1:i1 + 8]))
    a = 0.5 * (w[j0] + w[j0 + 1])
    b = 0.5 * (w[j1] + w[j1 + 1])
    return a, b


@pytest.mark.gpu
def test_assembled_operator_matches_same_operator(assembled):
    from normalmodes_amd import pevsl as pv
    from normalmodes_amd.driver import _open_solver
    name, Q, mv, S, B, w = assembled
    assert mv.fs.apdeg == (100 if Q.porder == 2 else 25)
    St = S * mv.d[:, None] * mv.d[None, :]
    rng = np.random.default_rng(5)
    x = rng.standard_normal(mv.n)
    h = _open_solver(mv, precision=pv.FP64, seed=7, rtol=1e-9)
    try:
        y = pv.apply(h, 0, x)
    finally:
        pv.pevsl_finish(h)
    ref = St @ x
    assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref), name


def _band(w, center, n_modes=16):
    """An interval of ~n_modes eigenvalues from the one nearest ``center``,
    its ends in the widest gaps nearby."""
    i0 = int(np.argmin(np.abs(w - center)))
    i1 = min(i0 + n_modes, w.size - 2)
    g = np.diff(w)
    j0 = i0 - 3 + int(np.argmax(g[i0 - 3:i0 + 3]))
    j1 = i1 + int(np.argmax(g[i1:i1 + 6]))
    return 0.5 * (w[j0] + w[j0 + 1]), 0.5 * (w[j1] + w[j1 + 1])


def _check_pairs(name, mv, S, B, r, ref, tol):
    assert r.lam.size == ref.size, (name, r.lam.size, ref.size)
    err = np.abs(r.lam - ref) / ref
    print(f"{name}: {ref.size} pairs, max rel err {err.max():.2e} (tolerance {tol.min():.1e}"
          f"..{tol.max():.1e})", flush=True)
    assert np.all(err <= tol), name
    assert np.all(r.res2 <= 1e-10) and r.target_met
    # recomputed residual ||S~x - lam B~x|| / |lam| of the scaled pencil
    St = S * mv.d[:, None] * mv.d[None, :]
    Bt = B * mv.d[:, None] * mv.d[None, :]
    X = r.X
    R = St @ X.T - (Bt @ X.T) * r.lam[None, :]
    xb = np.einsum("ij,ji->i", X, Bt @ X.T)
    r2 = np.linalg.norm(R, axis=0) / np.sqrt(xb) / np.abs(r.lam)
    assert np.all(r2 <= 1e-10), r2
    # two evaluations of the same residual differ by its own rounding, about
    # eps * lam_max / |lam| (the product S~x carries eps * ||S~|| ||x||)
    floor = 0.1 * np.finfo(float).eps * r.lmax / np.abs(r.lam)
    assert np.all(np.abs(r2 - r.res2) <= 1e-3 * r2 + np.maximum(floor, 1e-14))


@pytest.mark.gpu
def test_assembled_eigenpairs_match_same_operator(assembled):
    """All eigenpairs of a band at lam_max/1000, where the FP64 rounding floor
    eps lam_max / lam is ~2e-13: block ChebLanNr against
    the same operator's shift-invert eigenvalues at 1e-12 relative, and the
    residual gate recomputed from the returned vectors."""
    from normalmodes_amd.driver import apply_pevsl
    name, Q, mv, S, B, w = assembled
    a, b = _band(w, w[-1] / 1000.0)
    ref = orc.shift_invert_eigs(S, B, a, b)
    assert ref.size == w[(w >= a) & (w <= b)].size >= 10
    assert np.finfo(float).eps * w[-1] / ref.min() < 4e-13
    r = apply_pevsl(mv, interval=(a, b), block=4, verbose=20)
    _check_pairs(name + " mid-band", mv, S, B, r, ref, np.full(ref.size, 1e-12))


@pytest.mark.gpu
@pytest.mark.parametrize("block,split", [(2, 2), (4, 2), (4, 4)])
def test_assembled_split_chains_and_auto_ap_degree(assembled, block, split):
    """Filter chains of the fluid-solid operator on concurrent streams (one
    or two columns each) with the Ap~ solve at the degree of its own error
    bound (apdeg="auto", below the reference's 100): the same mid-band
    eigenpairs as the degree-100 operator's shift-invert reference, 1e-12."""
    from normalmodes_amd.driver import apply_pevsl, setupmatvec_fs
    name, Q, mv, S, B, w = assembled
    if name != "core-p2-job2":
        pytest.skip("one body suffices for the stream split")
    a, b = _band(w, w[-1] / 1000.0)
    ref = orc.shift_invert_eigs(S, B, a, b)
    mv2 = setupmatvec_fs(Q, bdeg="auto", apdeg="auto")
    try:
        assert mv2.fs.apdeg < mv.fs.apdeg == 100
        r = apply_pevsl(mv2, interval=(a, b), block=block, split=split, verbose=20)
    finally:
        mv2.free()
    assert r.lam.size == ref.size
    err = np.max(np.abs(r.lam - ref) / ref)
    print(f"{name} block {block} split {split} apdeg {mv2.fs.apdeg}: max rel err {err:.2e}", flush=True)
    assert err < 1e-12
    assert np.all(r.res2 <= 1e-10) and r.target_met


@pytest.mark.gpu
def test_assembled_refined_solves(assembled):
    """The fluid-solid operator with both Chebyshev solves (B~ and Ap~) in
    mixed-precision refinement: mid-band eigenpairs as the FP64 operator's."""
    from normalmodes_amd.driver import apply_pevsl, refine_bsolves
    name, Q, mv, S, B, w = assembled
    if name != "core-p2-job2":
        pytest.skip("one body suffices")
    a, b = _band(w, w[-1] / 1000.0)
    ref = orc.shift_invert_eigs(S, B, a, b)
    try:
        degs = refine_bsolves(mv, passes=2)
        r = apply_pevsl(mv, interval=(a, b), block=4, split=2, verbose=20)
    finally:
        refine_bsolves(mv, passes=0)
    assert r.lam.size == ref.size
    err = np.max(np.abs(r.lam - ref) / ref)
    print(f"{name} refined solves {degs}: max rel err {err:.2e}, max residual {r.res2.max():.2e}", flush=True)
    assert err < 1e-12
    assert np.all(r.res2 <= 1e-10) and r.target_met


@pytest.mark.gpu_long
def test_assembled_seismic_band(assembled):
    """The ~20 lowest seismic modes above the fluid's zero-energy cluster.
    There lam_max / lam is ~2e4 for these Earth-sized bodies, and every FP64
    evaluation of the operator (the device's products, the dense reference's
    formation of S and its factorisation) carries absolute eigenvalue errors
    of order eps lam_max: on the fluid-core body LAPACK's dense solver and a
    shift-invert reference differ by 1.3e-12 relative.  The bar is 1e-12
    relative or that floor, 8 eps lam_max / lam, whichever is larger."""
    from normalmodes_amd.driver import apply_pevsl
    name, Q, mv, S, B, w = assembled
    a, b = _seismic_interval(w)
    # dense LAPACK eigenvalues (their own error, ~eps lam_max, is far below
    # this band's tolerance; shift-invert converges slowly next to the
    # fluid's zero-energy cluster)
    ref = w[(w >= a) & (w <= b)]
    assert ref.size >= 10
    print(f"{name}: seismic band [{a:.6e}, {b:.6e}], {ref.size} eigenvalues", flush=True)
    r = apply_pevsl(mv, interval=(a, b), block=4, verbose=20)
    tol = np.maximum(1e-12, 8.0 * np.finfo(float).eps * w[-1] / ref)
    _check_pairs(name + " seismic band", mv, S, B, r, ref, tol)
    assert np.all(r.res2 <= 1e-10) and r.target_met
    # recomputed residual ||S~x - lam B~x|| / |lam| of the scaled pencil
    St = S * mv.d[:, None] * mv.d[None, :]
    Bt = B * mv.d[:, None] * mv.d[None, :]
    X = r.X
    R = St @ X.T - (Bt @ X.T) * r.lam[None, :]
    xb = np.einsum("ij,ji->i", X, Bt @ X.T)
    r2 = np.linalg.norm(R, axis=0) / np.sqrt(xb) / np.abs(r.lam)
    assert np.all(r2 <= 1e-10), r2
    # two evaluations of the same residual differ by its own rounding, about
    # eps * lam_max / |lam| (the product S~x carries eps * ||S~|| ||x||)
    floor = 0.1 * np.finfo(float).eps * r.lmax / np.abs(r.lam)
    assert np.all(np.abs(r2 - r.res2) <= 1e-3 * r2 + np.maximum(floor, 1e-14))
