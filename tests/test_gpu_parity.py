"""Parity of the HIP path (through the C-ABI) with the CPU oracle and with
dense FP64 eigenvalues, on seeded small meshes."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _pv():
    from normalmodes_amd import pevsl as pv
    return pv


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _recomputed_residuals(P, lam, X):
    """Both residual definitions of the scaled pencil recomputed with scipy
    from copied-out eigenvectors (rows of X, scaled coordinates x~):
    the reference's sqrt(sum r^2 / N)/|lam| (src/mod_pevsl.f90:137-162) and
    ||A~x - lam B~x||_2 / (|lam| sqrt(x^T B~ x))."""
    import scipy.sparse as sp
    A, B = P.A.to_scipy(), P.B.to_scipy()
    D = sp.diags(1.0 / np.sqrt(B.diagonal()))
    At, Bt = (D @ A @ D).tocsr(), (D @ B @ D).tocsr()
    R = (At @ X.T) - (Bt @ X.T) * lam[None, :]
    xb = np.einsum("ij,ji->i", X, Bt @ X.T)
    al = np.where(np.abs(lam) > 0, np.abs(lam), 1.0)
    nr = np.linalg.norm(R, axis=0)
    return nr / np.sqrt(P.n) / al, nr / np.sqrt(xb) / al


def _check_gate(P, r, gate=1e-10):
    """The device's nmg_residuals agree with the recomputation, and the
    recomputed 2-norm residual meets the gate."""
    r1, r2 = _recomputed_residuals(P, r.lam, r.X)
    assert np.all(r2 <= gate), r2
    # two evaluations of the same residual differ by its own rounding: the
    # product A~x carries eps ||A~|| ||x||, i.e. ~eps lam_max / |lam| relative
    # to |lam| (and never less than ~1e-14)
    floor = np.maximum(1e-14, 0.1 * np.finfo(float).eps * abs(r.lmax) / np.abs(r.lam))
    assert np.all(np.abs(r.res2 - r2) <= 1e-3 * r2 + floor), (r.res2, r2)
    assert np.all(np.abs(r.res_ref - r1) <= 1e-3 * r1 + floor / np.sqrt(P.n)), (r.res_ref, r1)


def test_spmv_layouts(small_p2):
    pv = _pv()
    P = small_p2
    rng = np.random.default_rng(1)
    x = rng.standard_normal(P.n)
    for M, layout in ((P.A, "bsr3"), (P.B, "mass3")):
        h = pv.csr_create(M.rowptr, M.col, M.val)
        try:
            assert pv.csr_info(h)["layout"] == layout
            y = pv.parcsr_matvec(x, h)
            ref = orc.spmv(M, x)
            scale = np.abs(M.to_scipy()) @ np.abs(x)
            assert np.max(np.abs(y - ref) / scale) < 1e-14
        finally:
            pv.csr_free(h)


def test_spmv_generic_csr():
    pv = _pv()
    import scipy.sparse as sp
    rng = np.random.default_rng(2)
    M = sp.random(777, 777, density=0.02, random_state=3, format="csr") + sp.eye(777, format="csr")
    M = M.tocsr()
    M.sort_indices()
    x = rng.standard_normal(777)
    h = pv.csr_create(M.indptr.astype(np.int64), M.indices, M.data)
    try:
        assert pv.csr_info(h)["layout"] == "csr"
        y = pv.parcsr_matvec(x, h)
        assert np.max(np.abs(y - M @ x)) < 1e-13 * np.max(np.abs(M) @ np.abs(x))
    finally:
        pv.csr_free(h)


def test_parcsrcreate_int32(small_p2):
    pv = _pv()
    B = small_p2.B
    n = B.n_rows
    h = pv.pevsl_parcsrcreate(n, n, [0, n], [0, n], B.rowptr.astype(np.int32), B.col, B.val)
    try:
        x = np.linspace(-1, 1, n)
        assert _rel(pv.parcsr_matvec(x, h), orc.spmv(B, x)) < 1e-14
    finally:
        pv.csr_free(h)


def test_diagscale(small_p2):
    pv = _pv()
    B = small_p2.B
    h = pv.csr_create(B.rowptr, B.col, B.val)
    try:
        d = pv.csr_diagscale(h, 1.0)
        d_ref, val_ref = orc.diagscale(B, 1.0)
        assert _rel(d, d_ref) < 1e-15
        x = np.random.default_rng(4).standard_normal(B.n_rows)
        import scipy.sparse as sp
        Bt = sp.csr_matrix((val_ref, B.col, B.rowptr), shape=(B.n_rows, B.n_rows))
        assert _rel(pv.parcsr_matvec(x, h), Bt @ x) < 1e-14
        # unit diagonal after scaling
        e = np.zeros(B.n_rows)
        e[5] = 1.0
        assert abs(pv.parcsr_matvec(e, h)[5] - 1.0) < 1e-15
    finally:
        pv.csr_free(h)


def test_diagscale_rejects_nonpositive():
    pv = _pv()
    import scipy.sparse as sp
    M = sp.diags([1.0, -2.0, 3.0]).tocsr()
    h = pv.csr_create(M.indptr.astype(np.int64), M.indices, M.data)
    try:
        with pytest.raises(pv.PevslError):
            pv.csr_diagscale(h, 1.0)
    finally:
        pv.csr_free(h)


def test_chebiter(small_p2):
    pv = _pv()
    B = small_p2.B
    h = pv.csr_create(B.rowptr, B.col, B.val)
    d_ref, val_ref = orc.diagscale(B, 1.0)
    try:
        pv.csr_diagscale(h, 1.0)
        b = np.random.default_rng(5).standard_normal(B.n_rows)
        for deg in (1, 2, 7, 45):
            c = pv.pevsl_setup_chebiter(0.28, 4.4, deg, h)
            x = pv.pevsl_chebiter(2, b, c)
            ref = orc.chebiter(orc.CSRView(B.rowptr, B.col, val_ref), 0.28, 4.4, deg, b)
            assert _rel(x, ref) < 1e-13, deg
            pv.pevsl_chebiter_free(c)
    finally:
        pv.csr_free(h)


def test_chebiter_persistent_matches_steps(small_p2, layered_p2):
    """The one-launch persistent solve against one launch per step (same
    recurrence, different summation order), FP64 and FP32 values, odd and
    even degrees; generic CSR matrices are not planned for it."""
    pv = _pv()
    for P in (small_p2, layered_p2):
        B = P.B
        h = pv.csr_create(B.rowptr, B.col, B.val)
        d_ref, val_ref = orc.diagscale(B, 1.0)
        try:
            pv.csr_diagscale(h, 1.0)
            b = np.random.default_rng(9).standard_normal(B.n_rows)
            for deg in (3, 58):
                c = pv.pevsl_setup_chebiter(0.25, 4.6, deg, h)
                try:
                    info = pv.chebiter_info(c)
                    assert not info["persistent"] and info["plannable"] and info["grid"] >= 1, info
                    pv.chebiter_set_mode(c, 1)
                    assert pv.chebiter_info(c)["persistent"]
                    x1 = pv.pevsl_chebiter(2, b, c)
                    pv.chebiter_set_mode(c, 0)
                    x0 = pv.pevsl_chebiter(2, b, c)
                    assert _rel(x1, x0) < 1e-14, deg
                    ref = orc.chebiter(orc.CSRView(B.rowptr, B.col, val_ref), 0.25, 4.6, deg, b)
                    assert _rel(x1, ref) < 1e-13, deg
                    pv.chebiter_set_mode(c, 1)
                    t1, _ = pv.bench_bsolve(c, 3)
                    assert t1 > 0
                    # K interleaved columns (block Lanczos, slices, KPM)
                    h2 = pv.pevsl_start(0)
                    try:
                        pv.pevsl_setprobsizes(h2, B.n_rows, B.n_rows, -1)
                        pv.pevsl_setbsol_chebiter(h2, 2, c)
                        for K in (2, 4, 16):
                            X = np.random.default_rng(K).standard_normal((K, B.n_rows))
                            Y1 = pv.apply_block(h2, 2, X)
                            pv.chebiter_set_mode(c, 0)
                            Y0 = pv.apply_block(h2, 2, X)
                            pv.chebiter_set_mode(c, 1)
                            assert _rel(Y1, Y0) < 1e-14, K
                            for q in (0, K - 1):
                                assert _rel(Y1[q], pv.pevsl_chebiter(2, X[q], c)) < 1e-14, (K, q)
                    finally:
                        pv.pevsl_finish(h2)
                finally:
                    pv.pevsl_chebiter_free(c)
        finally:
            pv.csr_free(h)
    A = small_p2.A
    g = pv.csr_create(A.rowptr, A.col, A.val)   # 3x3 blocks: not a scalar mass
    try:
        c = pv.pevsl_setup_chebiter(1.0, 2.0, 4, g)
        assert not pv.chebiter_info(c)["plannable"]
        pv.pevsl_chebiter_free(c)
    finally:
        pv.csr_free(g)


def test_findpol_matches_oracle():
    pv = _pv()
    for xintv in ([0.5, 12.0, -0.1, 12600.0], [3000.0, 3200.0, 0.0, 12600.0], [11000, 12600, 0, 12600]):
        p = pv.pevsl_findpol(xintv, 0.8, 0.7)
        info = pv.pol_info(p)
        ref = orc.findpol(xintv, 0.8, 0.7)
        assert info.deg == ref.deg
        assert abs(info.bar - ref.bar) < 1e-12
        assert np.max(np.abs(info.mu - ref.mu)) < 1e-12
        pv.pevsl_freepol(p)


@pytest.fixture(scope="module")
def setup_small(small_p2):
    from normalmodes_amd.driver import setupmatvec
    mv = setupmatvec(small_p2)
    yield mv
    mv.free()


def test_setup_bounds(small_p2, setup_small):
    mv = setup_small
    P = orc.Pencil(small_p2.A, small_p2.B, bdeg=45)
    assert abs(mv.blmin - P.blmin) < 1e-9 * P.blmax
    assert abs(mv.blmax - P.blmax) < 1e-9 * P.blmax
    import scipy.sparse as sp
    ev = np.linalg.eigvalsh(sp.csr_matrix((P.Bt.val, P.Bt.col, P.Bt.rowptr)).toarray())
    assert mv.blmin <= ev[0] * (1 + 1e-9) and mv.blmax >= ev[-1] * (1 - 1e-9)


def test_operators_and_filter(small_p2, setup_small):
    pv = _pv()
    mv = setup_small
    P = orc.Pencil(small_p2.A, small_p2.B, bdeg=mv.bdeg, bounds_B=(mv.blmin, mv.blmax))
    h = pv.pevsl_start(0)
    try:
        pv.pevsl_setprobsizes(h, mv.n, mv.n, -1)
        pv.pevsl_setbmv(h, mv.matB)
        pv.pevsl_setbsol_chebiter(h, 2, mv.chebB)
        pv.pevsl_setamv_dad(h, mv.matA, mv.d)
        pv.pevsl_set_geneig(h)
        x = np.random.default_rng(6).standard_normal(mv.n)
        assert _rel(pv.apply(h, 0, x), P.apply_a(x)) < 1e-13
        assert _rel(pv.apply(h, 1, x), P.apply_b(x)) < 1e-14
        assert _rel(pv.apply(h, 2, x), P.bsolve(x)) < 1e-13
        xintv = [0.5, 12.0, -0.1, 12600.0]
        pol = pv.pevsl_findpol(xintv, 0.8, 0.7)
        y = pv.filter_apply(h, pol, x)
        yr = P.chebav(orc.findpol(xintv, 0.8, 0.7), x)
        assert _rel(y, yr) < 1e-11
        pv.pevsl_freepol(pol)
    finally:
        pv.pevsl_finish(h)


def test_full_path_matches_dense(small_p2, setup_small, dense_eigs):
    from normalmodes_amd.driver import apply_pevsl
    mv = setup_small
    r = apply_pevsl(mv, interval=(1.0, 12.0))
    w = dense_eigs(small_p2)
    ref = w[(w >= 1.0) & (w <= 12.0)]
    assert r.lam.size == ref.size
    assert np.max(np.abs(r.lam - ref) / ref) < 1e-12
    # residual of the reference's definition is small; the 2-norm gate needs
    # the auto B-solve degree (test below)
    assert np.all(r.res_ref < 1e-9)
    # oracle restatement of the same path
    lam_o, X_o, info = orc.solve_interval(small_p2.A, small_p2.B, (1.0, 12.0), bdeg=mv.bdeg,
                                          return_info=True)
    assert lam_o.size == r.lam.size
    assert np.max(np.abs(r.lam - lam_o) / lam_o) < 1e-12
    assert abs(r.lmax - info["lmax"]) < 1e-3 * info["lmax"]
    assert r.deg == info["deg"] or abs(r.deg - info["deg"]) <= 2


def test_full_path_residual_gate(small_p2, dense_eigs):
    from normalmodes_amd.driver import setupmatvec, apply_pevsl
    mv = setupmatvec(small_p2, bdeg="auto")
    try:
        r = apply_pevsl(mv, interval=(1.0, 12.0))
        w = dense_eigs(small_p2)
        ref = w[(w >= 1.0) & (w <= 12.0)]
        assert r.lam.size == ref.size
        assert np.max(np.abs(r.lam - ref) / ref) < 1e-12
        assert np.all(r.res2 <= 1e-10), r.res2
        _check_gate(small_p2, r)
        # eigenvectors: B~-orthonormal
        X = r.X
        BX = np.stack([orc.Pencil(small_p2.A, small_p2.B, bdeg=1, bounds_B=(1, 2)).apply_b(x) for x in X])
        G = X @ BX.T
        assert np.max(np.abs(G - np.eye(len(X)))) < 1e-10
    finally:
        mv.free()


def test_from_lmin_includes_rigid_modes(small_p2, setup_small, dense_eigs):
    from normalmodes_amd.driver import apply_pevsl
    r = apply_pevsl(setup_small, interval=(0.0, 7.0))
    w = dense_eigs(small_p2)
    assert r.lam.size == np.sum(w <= 7.0)
    assert np.all(np.abs(r.lam[:6]) < 1e-8)
    assert np.max(np.abs(r.lam[6:] - w[6:r.lam.size]) / w[6:r.lam.size]) < 1e-12


def test_p1_reference_degree(small_p1, dense_eigs):
    from normalmodes_amd.driver import setupmatvec, apply_pevsl
    mv = setupmatvec(small_p1)
    try:
        assert mv.bdeg == 25
        r = apply_pevsl(mv, interval=(1.0, 14.0))
        w = dense_eigs(small_p1)
        ref = w[(w >= 1.0) & (w <= 14.0)]
        assert r.lam.size == ref.size
        assert np.max(np.abs(r.lam - ref) / ref) < 1e-12
        assert np.all(r.res2 <= 1e-10)
    finally:
        mv.free()


def test_bad_handle_and_state():
    pv = _pv()
    with pytest.raises(pv.PevslError):
        pv.pevsl_finish(12345)
    h = pv.pevsl_start(0)
    try:
        pv.pevsl_setprobsizes(h, 10, 10)
        with pytest.raises(pv.PevslError):
            pv.pevsl_lanbounds(h, 30, 30, 1e-5)  # no operator set
        with pytest.raises(pv.PevslError):
            pv.pevsl_set_geneig(h)               # no B
    finally:
        pv.pevsl_finish(h)


def test_c1_matches_shift_invert_golden():
    """The C1 workload (~52k DoF) on the device against the committed
    shift-invert eigenvalues: same count in the interval, 1e-12 relative."""
    import json
    from pathlib import Path
    from normalmodes_amd.driver import setupmatvec, apply_pevsl
    from normalmodes_amd.workloads import build_problem, INTERVALS
    g = json.loads((Path(__file__).resolve().parent / "golden" / "c1_eigs.json").read_text())
    w = np.array(g["eigenvalues"])
    a, b = INTERVALS["c1"]
    ref = w[(w >= a) & (w <= b)]
    P = build_problem("c1")
    assert P.n == g["n_dof"] and P.A.nnz == g["nnz_A"]
    mv = setupmatvec(P, bdeg="auto")
    try:
        r = apply_pevsl(mv, interval=(a, b))
    finally:
        mv.free()
    assert r.lam.size == ref.size == g["n_in_interval"]
    assert np.max(np.abs(r.lam - ref) / ref) < 1e-12
    assert np.all(r.res2 <= 1e-10) and r.target_met
    _check_gate(P, r)


@pytest.mark.parametrize("block", [2, 4, 8, 16, 32, 64])
def test_block_lanczos_matches_dense(small_p2, dense_eigs, block):
    """The block filtered Lanczos (same filter, block B-solve, GEMM
    re-orthogonalisation) finds the same eigenpairs as the dense solver."""
    from normalmodes_amd.driver import setupmatvec, apply_pevsl
    mv = setupmatvec(small_p2, bdeg="auto")
    try:
        r = apply_pevsl(mv, interval=(1.0, 12.0), block=block)
    finally:
        mv.free()
    w = dense_eigs(small_p2)
    ref = w[(w >= 1.0) & (w <= 12.0)]
    assert r.lam.size == ref.size
    assert np.max(np.abs(r.lam - ref) / ref) < 1e-12
    assert np.all(r.res2 <= 1e-10), r.res2
    _check_gate(small_p2, r)


def test_block_products_match_single_column(small_p2):
    """k-column sparse products equal k single-column products (A in BSR3,
    B in the scalar-mass layout, generic CSR)."""
    pv = _pv()
    from normalmodes_amd.driver import setupmatvec
    mv = setupmatvec(small_p2, bdeg="auto")
    h = pv.pevsl_start(0)
    try:
        pv.pevsl_setprobsizes(h, mv.n, mv.n, -1)
        pv.pevsl_setbmv(h, mv.matB)
        pv.pevsl_setbsol_chebiter(h, 2, mv.chebB)
        pv.pevsl_setamv_dad(h, mv.matA, mv.d)
        pv.pevsl_set_geneig(h)
        X = np.random.default_rng(9).standard_normal((8, mv.n))
        for which in (0, 1, 2):
            Y = pv.apply_block(h, which, X)
            ref = np.stack([pv.apply(h, which, x) for x in X])
            assert _rel(Y, ref) < 1e-13, which
    finally:
        pv.pevsl_finish(h)
        mv.free()


def test_fp32_bsolve_eigenvalues_match_fp64(small_p2, dense_eigs):
    """FP32 values in the Chebyshev B-solve: eigenvalues still agree with the
    FP64 path to 1e-12 relative (quadratic in the vector error); the 2-norm
    residual is bounded by the FP32 B-solve near 1e-6 and is reported."""
    from normalmodes_amd import pevsl as pv
    from normalmodes_amd.driver import setupmatvec, apply_pevsl
    mv = setupmatvec(small_p2, bdeg="auto")
    try:
        r64 = apply_pevsl(mv, interval=(1.0, 12.0))
        r32 = apply_pevsl(mv, interval=(1.0, 12.0), precision=pv.FP32_FILTER)
    finally:
        mv.free()
    assert r32.lam.size == r64.lam.size
    assert np.max(np.abs(r32.lam - r64.lam) / r64.lam) < 1e-12
    assert np.all(r32.res2 < 1e-4)


def test_refined_bsolve_accuracy(small_p2):
    """Mixed-precision refinement of the B~ solve: one FP32 pass stops near
    the FP32 floor, two passes reach ~1e-12, three the FP64 solve's level;
    measured against the exact solve of the same scaled mass (sparse LU),
    single column and 8 columns."""
    import scipy.sparse.linalg as spl
    pv = _pv()
    from normalmodes_amd.driver import setupmatvec, refine_inner_degree
    mv = setupmatvec(small_p2, bdeg="auto")
    h = pv.pevsl_start(0)
    try:
        pv.pevsl_setprobsizes(h, mv.n, mv.n, -1)
        pv.pevsl_setbmv(h, mv.matB)
        pv.pevsl_setbsol_chebiter(h, 2, mv.chebB)
        pv.pevsl_setamv_dad(h, mv.matA, mv.d)
        pv.pevsl_set_geneig(h)
        import scipy.sparse as sp
        Bt = (sp.diags(mv.d) @ small_p2.B.to_scipy() @ sp.diags(mv.d)).tocsc()
        lu = spl.splu(Bt)
        X = np.random.default_rng(11).standard_normal((8, mv.n))
        ref = np.stack([lu.solve(x) for x in X])
        deg = refine_inner_degree(mv.blmin, mv.blmax)
        errs = {}
        for passes in (0, 1, 2, 3):
            pv.chebiter_set_refine(mv.chebB, passes, deg)
            y1 = pv.apply(h, 2, X[0])
            Y = pv.apply_block(h, 2, X)
            errs[passes] = (_rel(y1, ref[0]), _rel(Y, ref))
        pv.chebiter_set_refine(mv.chebB, 0)
    finally:
        pv.pevsl_finish(h)
        mv.free()
    print("refined B-solve errors (1 column, 8 columns) by passes:", errs, flush=True)
    for k in (0, 1):
        assert errs[0][k] < 1e-13
        assert 1e-9 < errs[1][k] < 1e-5          # FP32 floor after one pass
        assert errs[2][k] < 1e-11
        assert errs[3][k] < 3e-13


@pytest.mark.parametrize("block", [1, 8])
def test_refined_bsolve_solve_meets_gate(small_p2, dense_eigs, block):
    """The whole solve with the refined B-solve: the same eigenvalues as the
    dense solver to 1e-12 and the 1e-10 residual gate, recomputed."""
    from normalmodes_amd.driver import setupmatvec, apply_pevsl, refine_bsolves
    mv = setupmatvec(small_p2, bdeg="auto")
    try:
        refine_bsolves(mv, passes=2)
        r = apply_pevsl(mv, interval=(1.0, 12.0), block=block)
    finally:
        mv.free()
    w = dense_eigs(small_p2)
    ref = w[(w >= 1.0) & (w <= 12.0)]
    assert r.lam.size == ref.size
    assert np.max(np.abs(r.lam - ref) / ref) < 1e-12
    assert np.all(r.res2 <= 1e-10) and r.target_met
    _check_gate(small_p2, r)


@pytest.mark.parametrize("block", [1, 8])
def test_fp32_stiffness_filter_error_stated(small_p2, dense_eigs, block):
    """FP32 values of A~ in the filter products as well (NMG_FP32_FILTER_A):
    the filtered subspace then belongs to the FP32-rounded pencil, whose low
    modes move by ~eps32 lam_max / lam relative to their gaps; Rayleigh-Ritz
    and the residual stay FP64, so the eigenvalues keep the error quadratic
    in that vector error.  The per-eigenvalue error against the FP64 path
    and the residual are printed and bounded, not gated at 1e-10."""
    from normalmodes_amd import pevsl as pv
    from normalmodes_amd.driver import setupmatvec, apply_pevsl
    mv = setupmatvec(small_p2, bdeg="auto")
    try:
        r64 = apply_pevsl(mv, interval=(1.0, 12.0), block=block)
        r32 = apply_pevsl(mv, interval=(1.0, 12.0), block=block, precision=pv.FP32_FILTER_A)
    finally:
        mv.free()
    assert r32.lam.size == r64.lam.size
    err = np.abs(r32.lam - r64.lam) / r64.lam
    print(f"FP32 stiffness filter, block {block}: max rel eig err {err.max():.2e}, "
          f"max residual {r32.res2.max():.2e}", flush=True)
    # measured on MI355X: 1.8e-13 and 2.5e-6 (block 1 and 8)
    assert err.max() < 1e-11
    assert np.all(r32.res2 < 1e-4)


@pytest.mark.parametrize("block", [1, 8])
def test_empty_interval_stops_early(small_p2, dense_eigs, block):
    """An interval inside a spectral gap: no pair, and the stop comes from the
    converged largest Ritz value, not from maxit (single-vector and block)."""
    from normalmodes_amd.driver import setupmatvec, apply_pevsl
    w = dense_eigs(small_p2)
    nz = w[w > 1.0]
    g = int(np.argmax(np.diff(nz[:30])))
    a, b = nz[g] + 0.3 * (nz[g + 1] - nz[g]), nz[g] + 0.7 * (nz[g + 1] - nz[g])
    mv = setupmatvec(small_p2, bdeg="auto")
    try:
        r = apply_pevsl(mv, interval=(a, b), vectors=False, block=block)
    finally:
        mv.free()
    assert r.lam.size == 0 and r.lan_steps < 200 * block


@pytest.mark.parametrize("block", [1, 8])
def test_interval_beside_one_eigenvalue_stops(small_p2, dense_eigs, block):
    """An interval that starts just above an isolated eigenvalue: that
    eigenvalue's filtered Ritz value may reach the bar and converge while its
    Rayleigh quotient lies outside the interval.  The solve must stop with
    no pair, not run on to maxit."""
    from normalmodes_amd.driver import setupmatvec, apply_pevsl
    w = dense_eigs(small_p2)
    nz = w[w > 1.0]
    g = int(np.argmax(np.diff(nz[:30])))
    gap = nz[g + 1] - nz[g]
    a, b = nz[g] + 0.01 * gap, nz[g] + 0.4 * gap
    mv = setupmatvec(small_p2, bdeg="auto")
    try:
        r = apply_pevsl(mv, interval=(a, b), vectors=False, block=block)
    finally:
        mv.free()
    assert r.lam.size == 0 and r.lan_steps < 200 * block and r.target_met


def test_tall_skinny_products():
    """The split-reduction A^T B of the block re-orthogonalisation and the
    Rayleigh-Ritz Gram matrices against numpy, for every column width and
    odd sizes."""
    pv = _pv()
    rng = np.random.default_rng(12)
    for n, m, k in ((1, 1, 1), (1000, 3, 5), (51540, 424, 8), (20001, 37, 16), (777, 129, 37)):
        A = rng.standard_normal((n, m))
        B = rng.standard_normal((n, k))
        C = pv.gemm_tn(A, B)
        ref = A.T @ B
        assert np.max(np.abs(C - ref)) <= 1e-12 * np.sqrt(n) * np.abs(ref).max() + 1e-14, (n, m, k)


def test_tiled_products_match_untiled(small_p2):
    """LDS-tiled k-column products (scalar mass and 3x3 blocks; store and
    Chebyshev epilogues) for every width, rows per wavefront, tile capacity
    and workgroup size, against the untiled kernels."""
    from normalmodes_amd.driver import setupmatvec, _open_solver
    pv = _pv()
    mv = setupmatvec(small_p2, bdeg=20, tune=())
    M = small_p2.B.to_scipy()[0::3, 0::3].tocsr()
    longest = int(np.max(np.diff(M.indptr)))
    rng = np.random.default_rng(21)

    def products(X):
        # the solver clones A (scaled) with the layout A has when it is set
        h = _open_solver(mv, precision=pv.FP64, seed=7, rtol=1e-9)
        try:
            return [pv.apply_block(h, which, X) for which in (0, 1, 2)]
        finally:
            pv.pevsl_finish(h)

    try:
        for K in (1, 2, 4, 8, 16):
            X = rng.standard_normal((K, small_p2.n))
            for mat in (mv.matA, mv.matB):
                pv.set_tiling(mat, K, 0, 1)
            ref = products(X)
            limit = 64 * 1024 // (24 * K)
            for R in (1, 2, 4):
                if K * R > 64:
                    continue
                for cap in sorted({longest, min(2 * longest, limit), limit}):
                    for threads in (256, 1024):
                        for mat in (mv.matA, mv.matB):
                            pv.set_tiling(mat, K, cap, R, threads)
                        for which, (Y, Yr) in enumerate(zip(products(X), ref)):
                            assert _rel(Y, Yr) < 1e-14, (K, R, cap, threads, which)
            for mat in (mv.matA, mv.matB):
                pv.set_tiling(mat, K, 0, 1)
    finally:
        mv.free()


def test_bsr3_value_order():
    """3x3-block values in 64-entry tiles give the same products (single and
    k-column, untiled and LDS-tiled) and the same diagonal scaling as the
    plane order, bit for bit."""
    from normalmodes_amd.workloads import build_problem
    pv = _pv()
    P = build_problem("tiny")
    A = P.A
    rng = np.random.default_rng(31)
    x = rng.standard_normal(A.n_rows)
    X = rng.standard_normal((8, A.n_rows))
    ha = pv.csr_create(A.rowptr, A.col, A.val)
    hb = pv.csr_create(A.rowptr, A.col, A.val)
    try:
        assert pv.csr_info(ha)["layout"] == "bsr3"
        pv.set_value_order(hb, True)
        assert np.array_equal(pv.parcsr_matvec(x, ha), pv.parcsr_matvec(x, hb))
        da = pv.csr_diagscale(ha, 1.0)
        db = pv.csr_diagscale(hb, 1.0)
        assert np.array_equal(da, db)
        assert np.array_equal(pv.parcsr_matvec(x, ha), pv.parcsr_matvec(x, hb))
        for cap in (0, 120):
            for K in (1, 8):
                for h in (ha, hb):
                    pv.set_tiling(h, K, cap, 1)
                outs = []
                for h in (ha, hb):
                    s = pv.pevsl_start(0)
                    try:
                        pv.pevsl_setprobsizes(s, A.n_rows, A.n_rows, -1)
                        pv.pevsl_setamv(s, h)
                        outs.append(pv.apply_block(s, 0, X[:K]))
                    finally:
                        pv.pevsl_finish(s)
                assert np.array_equal(outs[0], outs[1]), (K, cap)
        pv.set_value_order(hb, False)
        assert np.array_equal(pv.parcsr_matvec(x, ha), pv.parcsr_matvec(x, hb))
    finally:
        pv.csr_free(ha)
        pv.csr_free(hb)


def test_split_stream_bsolve_and_solve(small_p2, dense_eigs):
    """k-column B-solves split over 2 and 4 streams: the same block products
    (to rounding) and the same eigenpairs as one stream."""
    from normalmodes_amd.driver import setupmatvec, apply_pevsl, _open_solver
    pv = _pv()
    mv = setupmatvec(small_p2, bdeg="auto", tune=(1, 2, 4, 8))
    try:
        X = np.random.default_rng(41).standard_normal((8, small_p2.n))
        out = {}
        for parts in (0, 2, 4):
            h = _open_solver(mv, precision=pv.FP64, seed=7, rtol=1e-9, split=parts)
            try:
                out[parts] = pv.apply_block(h, 2, X)
            finally:
                pv.pevsl_finish(h)
        assert _rel(out[2], out[0]) < 1e-14 and _rel(out[4], out[0]) < 1e-14
        w = dense_eigs(small_p2)
        ref = w[(w >= 1.0) & (w <= 12.0)]
        for parts in (2, 4):
            r = apply_pevsl(mv, interval=(1.0, 12.0), block=8, split=parts, vectors=False)
            assert r.lam.size == ref.size and np.max(np.abs(r.lam - ref) / ref) < 1e-12
            assert np.all(r.res2 <= 1e-10)
    finally:
        mv.free()


@pytest.mark.gpu
def test_graph_replayed_bsolves(small_p2, dense_eigs):
    """B-solves replayed from captured HIP graphs: four calls in a row on
    different inputs (eager first use, capture on the second, replays after),
    each against the same call without graphs, for one stream and for the
    2-stream split, and again after the B products are given new tiles; then
    whole solves with graphs, single-vector and block 8."""
    from normalmodes_amd.driver import setupmatvec, apply_pevsl, _open_solver
    pv = _pv()
    mv = setupmatvec(small_p2, bdeg="auto", tune=(1, 4, 8))
    try:
        rng = np.random.default_rng(43)
        Xs = [rng.standard_normal((8, small_p2.n)) for _ in range(4)]
        for parts in (0, 2):
            eager = _open_solver(mv, precision=pv.FP64, seed=7, rtol=1e-9, split=parts)
            graph = _open_solver(mv, precision=pv.FP64, seed=7, rtol=1e-9, split=parts, graphs=True)
            try:
                for X in Xs:
                    ref = pv.apply_block(eager, 2, X)
                    got = pv.apply_block(graph, 2, X)
                    assert _rel(got, ref) < 1e-15, parts
                # new LDS tiles for the B products of the widths in use (8
                # columns, 4 per part when split): the graphs captured above
                # hold the old kernels and tile buffers and must not replay
                longest = int(np.diff(small_p2.B.rowptr).max())
                for K in (4, 8):
                    pv.set_tiling(mv.matB, K, 2 * longest, 1)
                for X in Xs[:3]:
                    ref = pv.apply_block(eager, 2, X)
                    got = pv.apply_block(graph, 2, X)
                    assert _rel(got, ref) < 1e-15, parts
            finally:
                pv.pevsl_finish(eager)
                pv.pevsl_finish(graph)
        w = dense_eigs(small_p2)
        ref = w[(w >= 1.0) & (w <= 12.0)]
        for block in (1, 8):
            r0 = apply_pevsl(mv, interval=(1.0, 12.0), block=block, vectors=False)
            r = apply_pevsl(mv, interval=(1.0, 12.0), block=block, vectors=False, graphs=True)
            assert r.lam.size == ref.size and np.max(np.abs(r.lam - ref) / ref) < 1e-12
            assert np.max(np.abs(r.lam - r0.lam) / r0.lam) < 1e-14
            assert np.all(r.res2 <= 1e-10)
    finally:
        mv.free()


def _ragged_node_pattern(nb, seed):
    """Node pattern with empty rows, short rows and rows longer than 64 and
    128 entries (sorted unique columns), as a scipy CSR of ones."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    lens = rng.integers(1, 40, nb)
    lens[rng.choice(nb, 8, replace=False)] = 0
    lens[rng.choice(nb, 4, replace=False)] = [70, 130, 200, 65]
    rows, cols = [], []
    for r, L in enumerate(lens):
        c = np.sort(rng.choice(nb, size=min(int(L), nb), replace=False))
        rows.extend([r] * c.size)
        cols.extend(c.tolist())
    return sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(nb, nb))


def test_pipelined_row_products(small_p2, dense_eigs):
    """Single-column 3x3-block products through pipelined lane groups (16,
    32 and 64 lanes walking 2, 4 and 8 rows) on the P2 stiffness and on a
    ragged pattern with empty rows and rows of 65 to 200 blocks, against the
    CPU product; then a single-vector solve whose stiffness products and
    filter steps run on them, against dense eigenvalues."""
    import scipy.sparse as sp
    from normalmodes_amd.driver import setupmatvec, apply_pevsl
    pv = _pv()
    rng = np.random.default_rng(29)
    S = _ragged_node_pattern(300, 7)
    Sb = S.tocoo()
    A_r = sp.bsr_matrix((rng.standard_normal((Sb.nnz, 3, 3)), Sb.col, S.indptr), shape=(900, 900)).tocsr()
    A_r.sort_indices()
    for M in (small_p2.A.to_scipy(), A_r):
        h = pv.csr_create(M.indptr.astype(np.int64), M.indices.astype(np.int32), M.data)
        try:
            assert pv.csr_info(h)["layout"] == "bsr3"
            x = rng.standard_normal(M.shape[0])
            ref = M @ x
            scale = np.abs(M) @ np.abs(x) + 1e-300
            for lanes in (16, 32, 64):
                for rows in (2, 4, 8):
                    pv.set_row_pipeline(h, lanes, rows)
                    y = pv.parcsr_matvec(x, h)
                    assert np.max(np.abs(y - ref) / scale) < 1e-14, (lanes, rows)
            pv.set_row_pipeline(h, 0, 0)
            assert np.max(np.abs(pv.parcsr_matvec(x, h) - ref) / scale) < 1e-14
        finally:
            pv.csr_free(h)
    mv = setupmatvec(small_p2, bdeg="auto", tune=())
    try:
        pv.set_row_pipeline(mv.matA, 32, 4)
        r = apply_pevsl(mv, interval=(1.0, 12.0), vectors=False)
        w = dense_eigs(small_p2)
        ref = w[(w >= 1.0) & (w <= 12.0)]
        assert r.lam.size == ref.size and np.max(np.abs(r.lam - ref) / ref) < 1e-12
        assert np.all(r.res2 <= 1e-10)
        with pytest.raises(pv.PevslError):
            pv.set_row_pipeline(mv.matB, 32, 4)   # scalar mass: not 3x3 blocks
    finally:
        mv.free()


def test_row_group_products(small_p2, dense_eigs):
    """Single-column products through row groups (1, 2 and 8 groups per
    wavefront) for the stiffness (3x3 blocks) and mass (scalar x I3) of a P2
    mesh and for ragged synthetic patterns with empty rows and rows of 65 to
    200 entries, against the CPU product; Chebyshev B-solve and a whole solve
    with both matrices on row groups."""
    import scipy.sparse as sp
    pv = _pv()
    rng = np.random.default_rng(17)
    S = _ragged_node_pattern(300, 5)
    Sb = S.tocoo()
    blocks = rng.standard_normal((Sb.nnz, 3, 3))
    A_r = sp.bsr_matrix((blocks, Sb.col, S.indptr), shape=(900, 900)).tocsr()
    A_r.sort_indices()
    M_r = sp.kron(S.multiply(rng.uniform(0.5, 1.5, S.shape)).tocsr(), sp.eye(3), format="csr")
    M_r.sort_indices()
    mats = [(small_p2.A.to_scipy(), "bsr3"), (small_p2.B.to_scipy(), "mass3"), (A_r, "bsr3"), (M_r, "mass3")]
    for M, layout in mats:
        h = pv.csr_create(M.indptr.astype(np.int64), M.indices.astype(np.int32), M.data)
        try:
            assert pv.csr_info(h)["layout"] == layout
            x = rng.standard_normal(M.shape[0])
            ref = M @ x
            scale = np.abs(M) @ np.abs(x) + 1e-300
            for pw in (1, 2, 8):
                pv.set_row_groups(h, pw)
                y = pv.parcsr_matvec(x, h)
                assert np.max(np.abs(y - ref) / scale) < 1e-14, (layout, pw)
            pv.set_row_groups(h, 0)
            assert np.max(np.abs(pv.parcsr_matvec(x, h) - ref) / scale) < 1e-14
        finally:
            pv.csr_free(h)
    # Chebyshev B-solve and a whole solve on row groups
    B = small_p2.B
    d_ref, val_ref = orc.diagscale(B, 1.0)
    h = pv.csr_create(B.rowptr, B.col, B.val)
    try:
        pv.csr_diagscale(h, 1.0)
        pv.set_row_groups(h, 4)
        b = rng.standard_normal(B.n_rows)
        c = pv.pevsl_setup_chebiter(0.28, 4.4, 45, h)
        x = pv.pevsl_chebiter(2, b, c)
        ref = orc.chebiter(orc.CSRView(B.rowptr, B.col, val_ref), 0.28, 4.4, 45, b)
        assert _rel(x, ref) < 1e-13
        pv.pevsl_chebiter_free(c)
    finally:
        pv.csr_free(h)
    from normalmodes_amd.driver import setupmatvec, apply_pevsl
    mv = setupmatvec(small_p2, bdeg="auto", tune=())
    try:
        for mat in (mv.matA, mv.matB):
            pv.set_row_groups(mat, 2)
        w = dense_eigs(small_p2)
        ref = w[(w >= 1.0) & (w <= 12.0)]
        r = apply_pevsl(mv, interval=(1.0, 12.0), vectors=False)
        assert r.lam.size == ref.size and np.max(np.abs(r.lam - ref) / ref) < 1e-12
        assert np.all(r.res2 <= 1e-10)
    finally:
        mv.free()


def test_fp32_vector_bsolve(small_p2, dense_eigs):
    """FP32_BSOLVE: the k-column B-solve with FP32 values and FP32 iterates
    (FP64 product and update per entry) against the FP64 solve, one stream
    and split over two; then block-8 solves whose eigenvalues stay within
    1e-10 of the dense pencil's."""
    from normalmodes_amd.driver import setupmatvec, apply_pevsl, _open_solver
    pv = _pv()
    mv = setupmatvec(small_p2, bdeg="auto", tune=(1, 4, 8))
    try:
        X = np.random.default_rng(47).standard_normal((8, small_p2.n))
        h64 = _open_solver(mv, precision=pv.FP64, seed=7, rtol=1e-9, split=0)
        try:
            ref = pv.apply_block(h64, 2, X)
        finally:
            pv.pevsl_finish(h64)
        for parts in (0, 2):
            h = _open_solver(mv, precision=pv.FP32_BSOLVE, seed=7, rtol=1e-9, split=parts)
            try:
                for _ in range(2):
                    y = pv.apply_block(h, 2, X)
                    assert 1e-12 < _rel(y, ref) < 1e-5, (parts, _rel(y, ref))
            finally:
                pv.pevsl_finish(h)
        w = dense_eigs(small_p2)
        lam = w[(w >= 1.0) & (w <= 12.0)]
        for split in (0, 2):
            r = apply_pevsl(mv, interval=(1.0, 12.0), block=8, split=split, precision=pv.FP32_BSOLVE,
                            vectors=False)
            assert r.lam.size == lam.size and np.max(np.abs(r.lam - lam) / lam) < 1e-10
            assert np.all(r.res2 <= 1e-4)
    finally:
        mv.free()


def test_wide_mass_products(small_p2):
    """The wide k-column scalar-mass layout (one entry per lane, 4 to 32
    lanes per row), plain and on LDS tiles, gives the products, B-solves and
    FP32-vector B-solves of the untiled layout, for K = 2, 4, 8, 16."""
    from normalmodes_amd.driver import setupmatvec, _open_solver
    pv = _pv()
    mv = setupmatvec(small_p2, bdeg="auto", tune=())
    rng = np.random.default_rng(53)

    def products(X, prec):
        h = _open_solver(mv, precision=prec, seed=7, rtol=1e-9, split=0)
        try:
            return [pv.apply_block(h, which, X) for which in (1, 2)]
        finally:
            pv.pevsl_finish(h)

    try:
        for K in (2, 4, 8, 16):
            X = rng.standard_normal((K, small_p2.n))
            pv.set_tiling(mv.matB, K, 0, 1)
            ref = products(X, pv.FP64)
            ref32 = products(X, pv.FP32_BSOLVE)
            for G in (4, 8, 16, 32):
                pv.set_wide(mv.matB, K, G)
                for which, (Y, Yr) in enumerate(zip(products(X, pv.FP64), ref)):
                    assert _rel(Y, Yr) < 1e-14, (K, G, which)
                Y32 = products(X, pv.FP32_BSOLVE)
                assert _rel(Y32[0], ref32[0]) < 1e-14, (K, G)
                assert _rel(Y32[1], ref32[1]) < 1e-6, (K, G)
            pv.set_wide(mv.matB, K, 0)
            # the same on LDS tiles (two tile sizes)
            longest = int(np.max(np.diff(small_p2.B.to_scipy().indptr)))
            for cap in (longest, min(3 * longest, 64 * 1024 // (24 * K))):
                for G in (4, 8, 16):
                    pv.set_tiled_wide(mv.matB, K, cap, G)
                    for which, (Y, Yr) in enumerate(zip(products(X, pv.FP64), ref)):
                        assert _rel(Y, Yr) < 1e-14, ("tiled", K, cap, G, which)
                    Y32 = products(X, pv.FP32_BSOLVE)
                    assert _rel(Y32[1], ref32[1]) < 1e-6, ("tiled", K, cap, G)
            pv.set_tiled_wide(mv.matB, K, 0, 0)
        with pytest.raises(pv.PevslError):
            pv.set_wide(mv.matA, 4, 8)   # 3x3 blocks: not a scalar mass
        with pytest.raises(pv.PevslError):
            pv.set_tiled_wide(mv.matA, 4, 64, 8)
    finally:
        mv.free()


def test_split_filter_chains(small_p2, dense_eigs):
    """Block filters run as independent per-stream chains of k/2 columns
    (the default with two streams) against B-solve-only splitting and one
    stream: the same eigenvalues, FP64 and FP32 values; and spectrum slices
    through the same path."""
    from normalmodes_amd.driver import setupmatvec, apply_pevsl, apply_slices, _open_solver
    pv = _pv()
    mv = setupmatvec(small_p2, bdeg="auto", tune=(1, 4, 8))
    try:
        w = dense_eigs(small_p2)
        ref = w[(w >= 1.0) & (w <= 12.0)]
        r0 = apply_pevsl(mv, interval=(1.0, 12.0), block=8, split=0, vectors=False)
        r1 = apply_pevsl(mv, interval=(1.0, 12.0), block=8, split=2, vectors=False)
        assert r1.lam.size == ref.size and np.max(np.abs(r1.lam - ref) / ref) < 1e-12
        assert np.max(np.abs(r1.lam - r0.lam) / r0.lam) < 1e-13 and np.all(r1.res2 <= 1e-10)
        r2 = apply_pevsl(mv, interval=(1.0, 12.0), block=8, split=2, vectors=False, precision=pv.FP32_FILTER)
        assert r2.lam.size == ref.size and np.max(np.abs(r2.lam - ref) / ref) < 1e-12
        # B-solve-only splitting through the switch gives the same pairs
        h = _open_solver(mv, precision=pv.FP64, seed=7, rtol=1e-9, block=8, split=2)
        try:
            pv.set_split_filter(h, False)
            lmin, lmax = pv.pevsl_lanbounds(h, 3000, 5000, 1e-5)
            pol = pv.pevsl_findpol([1.0, 12.0, lmin, lmax], 0.8, 0.7)
            pv.pevsl_cheblannr(h, [1.0, 12.0, lmin, lmax], 9624, 1e-5, pol)
            lam, _ = pv.pevsl_copy_result(h, small_p2.n, False)
            pv.pevsl_freepol(pol)
        finally:
            pv.pevsl_finish(h)
        lam = np.sort(lam)
        assert lam.size == ref.size and np.max(np.abs(lam - ref) / ref) < 1e-12
        # five slices run at width 8, so the slice filter takes the chains too
        # (wide band: narrow slices low in the spectrum need very high degrees)
        ref60 = w[(w >= 1.0) & (w <= 60.0)]
        s = apply_slices(mv, interval=(1.0, 60.0), n_slices=5, slicer="kpm", kpm_nvec=16, vectors=False)
        assert s.lam.size == ref60.size and np.max(np.abs(s.lam - ref60) / ref60) < 1e-12
    finally:
        mv.free()
