"""Pins of the CPU oracle (oracle/nm_oracle.c) against what does not depend
on it: scipy/numpy routines, closed forms, invariants and committed fixtures.
Each check is chosen so that a plausible slip (off-by-one degree, dropped
term, wrong sign or transposed operand) fails it; several tests also show
that a deliberately wrong variant is rejected."""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

from oracle import oracle as orc

GOLDEN = Path(__file__).resolve().parent / "golden"


def _rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)


# ---------------------------------------------------------------- dense helpers
def test_tridiag_eig_vs_scipy():
    rng = np.random.default_rng(0)
    for m in (1, 2, 7, 60, 200):
        d = rng.standard_normal(m)
        e = rng.standard_normal(max(m - 1, 0))
        w_ref, S_ref = sl.eigh_tridiagonal(d, e)
        w, S = orc.tridiag_eig(d, e)
        assert _rel(w, w_ref) < 1e-13
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        assert np.max(np.abs(T @ S - S * w[None, :])) < 1e-12 * max(1, np.abs(w).max())
        assert np.max(np.abs(S.T @ S - np.eye(m))) < 1e-12
        assert _rel(orc.tridiag_eigvals(d, e), w_ref) < 1e-13


def test_tridiag_clustered_and_graded():
    # Lanczos-like T with a tight cluster and widely graded entries
    m = 80
    d = np.concatenate([np.full(10, 1.0) + 1e-10 * np.arange(10), np.geomspace(1e-3, 1e6, m - 10)])
    e = np.full(m - 1, 1e-6)
    w_ref = sl.eigh_tridiagonal(d, e, eigvals_only=True)
    w, _ = orc.tridiag_eig(d, e)
    assert np.max(np.abs(w - w_ref)) < 1e-9


def test_sygv_vs_scipy():
    rng = np.random.default_rng(1)
    n = 40
    X = rng.standard_normal((n, n))
    GA = X + X.T
    Y = rng.standard_normal((n, n))
    GB = Y @ Y.T + n * np.eye(n)
    mu_ref = sl.eigh(GA, GB, eigvals_only=True)
    mu, W = orc.sygv(GA, GB)
    assert _rel(mu, mu_ref) < 1e-12
    assert np.max(np.abs(W.T @ GB @ W - np.eye(n))) < 1e-11
    assert np.max(np.abs(GA @ W - GB @ W * mu[None, :])) < 1e-10 * np.abs(GA).max()
    # a transposed-operand slip (G_B^T G_A) would not reproduce the pencil
    bad = sl.eigh(GA, GB @ GB.T, eigvals_only=True)
    assert _rel(mu, bad) > 1e-3


# ------------------------------------------------------------ sparse operators
def test_spmv_and_diagscale(small_p2):
    A, B = small_p2.A, small_p2.B
    x = np.random.default_rng(2).standard_normal(A.n_rows)
    assert _rel(orc.spmv(A, x), A.to_scipy() @ x) < 1e-14
    d, val = orc.diagscale(B, 1.0)
    Bs = B.to_scipy()
    assert _rel(d, 1.0 / np.sqrt(Bs.diagonal())) < 1e-15
    Bt = sp.csr_matrix((val, B.col, B.rowptr), shape=Bs.shape)
    D = sp.diags(d)
    assert abs(Bt - D @ Bs @ D).max() < 1e-15 * abs(Bs).max() / Bs.diagonal().min()
    assert np.max(np.abs(Bt.diagonal() - 1.0)) < 1e-15
    # sign = -1 (the Ap convention) on a negative-definite matrix
    d2, val2 = orc.diagscale(sp.csr_matrix(-Bs), -1.0)
    assert _rel(d2, d) < 1e-15 and _rel(val2, val) < 1e-15
    with pytest.raises(ValueError):
        orc.diagscale(sp.csr_matrix(-Bs), 1.0)


def test_dad_operator(small_p2):
    P = orc.Pencil(small_p2.A, small_p2.B, bdeg=5, bounds_B=(0.2, 5.0))
    x = np.random.default_rng(3).standard_normal(P.n)
    D = sp.diags(P.d)
    ref = D @ (small_p2.A.to_scipy() @ (D @ x))
    assert _rel(P.apply_a(x), ref) < 1e-14


# ----------------------------------------------------------------- chebiter
def _cheb_T(k, t):
    return np.cos(k * np.arccos(np.clip(t, -1, 1))) if np.all(np.abs(t) <= 1) else np.polynomial.chebyshev.chebval(
        t, np.eye(k + 1)[k])


def test_chebiter_closed_form(small_p2):
    """x = q(B) b with 1 - t q(t) = T_{c+1}((th - t)/de)/T_{c+1}(th/de): the
    Chebyshev semi-iteration after c products with B (x_{c+1})."""
    d, val = orc.diagscale(small_p2.B, 1.0)
    Bt = sp.csr_matrix((val, small_p2.B.col, small_p2.B.rowptr)).toarray()
    lam, U = np.linalg.eigh(Bt)
    lmin, lmax = 0.8 * lam[0], 1.05 * lam[-1]
    th, de = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    b = np.random.default_rng(4).standard_normal(Bt.shape[0])
    cb = U.T @ b
    for c in (0, 1, 2, 5, 30):
        x = orc.chebiter(sp.csr_matrix(Bt), lmin, lmax, c, b)
        k = c + 1
        T = np.polynomial.chebyshev.chebval
        R = T((th - lam) / de, np.eye(k + 1)[k]) / T(th / de, np.eye(k + 1)[k])
        ref = U @ ((1.0 - R) / lam * cb)
        assert _rel(x, ref) < 1e-12, c
        if 1 <= c <= 5:  # an off-by-one in the degree is detected (before convergence)
            Rw = T((th - lam) / de, np.eye(k)[k - 1]) / T(th / de, np.eye(k)[k - 1])
            assert _rel(x, U @ ((1.0 - Rw) / lam * cb)) > 1e-6
    # error bound 2 rho^(c+1) relative in the B-norm sense
    x = orc.chebiter(sp.csr_matrix(Bt), lmin, lmax, 40, b)
    xe = np.linalg.solve(Bt, b)
    kap = lmax / lmin
    rho = (np.sqrt(kap) - 1) / (np.sqrt(kap) + 1)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) <= 2 * rho ** 41 * kap


# ------------------------------------------------------------------ filter
def test_findpol_properties():
    T = np.polynomial.chebyshev.chebval
    for xintv, typ in (([0.5, 12.0, -0.1, 12600.0], 1), ([3000.0, 3200.0, 0.0, 12600.0], 0),
                       ([11000.0, 12600.0, 0.0, 12600.0], 2), ([3.14, 24.5, -0.02, 1.19e6], 1)):
        pol = orc.findpol(xintv, 0.8, 0.7)
        assert pol.type == typ
        thr = 0.8 if typ == 0 else 0.7
        cc, dd = 0.5 * (xintv[3] + xintv[2]), 0.5 * (xintv[3] - xintv[2])
        assert abs(pol.cc - cc) < 1e-12 * abs(cc) + 1e-300 and abs(pol.dd - dd) < 1e-12 * dd
        aa, bb = (max(xintv[0], xintv[2]) - cc) / dd, (min(xintv[1], xintv[3]) - cc) / dd
        assert abs(T(pol.gam, pol.mu) - 1.0) < 1e-12                   # normalised peak
        va, vb = T(aa, pol.mu), T(bb, pol.mu)
        if typ == 0:
            assert abs(va - vb) < 1e-10                                 # balanced
            assert max(va, vb) <= thr and abs(pol.bar - min(va, vb)) < 1e-12
            assert aa < pol.gam < bb
        elif typ == 1:
            assert vb <= thr and abs(pol.bar - vb) < 1e-12 and pol.gam == -1.0
        else:
            assert va <= thr and abs(pol.bar - va) < 1e-12 and pol.gam == 1.0
        # minimal degree: capping the search one below fails
        with pytest.raises(ValueError):
            orc.findpol(xintv, 0.8, 0.7, max_deg=pol.deg - 1)
        # the filter separates: larger inside the interval than far outside
        t_in = np.linspace(aa, bb, 50)[1:-1]
        t_out = np.linspace(min(bb + 0.5 * (1 - bb), 1.0), 1.0, 20) if typ != 2 else np.linspace(-1, aa - 0.5 * (aa + 1), 20)
        assert T(t_in, pol.mu).min() >= pol.bar - 1e-12
        assert np.abs(T(t_out, pol.mu)).max() < pol.bar


def test_lanczos_sigma_damping():
    """With damping 2 and a delta at t = 1 the coefficients are
    mu_k ~ g_k with g_0 = 1/2, g_k = sinc(k pi/(m+1))."""
    pol = orc.findpol([11000.0, 12600.0, 0.0, 12600.0], 0.8, 0.7)
    m = pol.deg
    k = np.arange(1, m + 1)
    g = np.concatenate([[0.5], np.sin(k * np.pi / (m + 1)) / (k * np.pi / (m + 1))])
    assert _rel(pol.mu / pol.mu[0], g / g[0]) < 1e-12


def test_chebav_standard_matches_chebval():
    """Standard problem with A = diag(l): ChebAv(e) = p((l - cc)/dd) e."""
    lam = np.linspace(0.0, 100.0, 257)
    A = sp.diags(lam).tocsr()
    P = orc.Pencil(A)
    pol = orc.findpol([10.0, 20.0, 0.0, 100.0], 0.8, 0.7)
    y = P.chebav(pol, np.ones_like(lam))
    ref = np.polynomial.chebyshev.chebval((lam - pol.cc) / pol.dd, pol.mu)
    assert _rel(y, ref) < 1e-12


def test_chebav_generalized_dense(small_p2):
    """y = B~ U p(Lambda) U^T z for the B~-orthonormal eigenvectors U (with an
    accurate B-solve)."""
    Pn = orc.Pencil(small_p2.A, small_p2.B, bdeg=200)
    At = sp.diags(Pn.d) @ small_p2.A.to_scipy() @ sp.diags(Pn.d)
    Bt = sp.csr_matrix((Pn.Bt.val, Pn.Bt.col, Pn.Bt.rowptr))
    w, U = sl.eigh(At.toarray(), Bt.toarray())
    xintv = [1.0, 12.0, w[0] - 1.0, w[-1] * 1.01]
    pol = orc.findpol(xintv, 0.8, 0.7)
    z = np.random.default_rng(5).standard_normal(Pn.n)
    y = Pn.chebav(pol, z)
    p = np.polynomial.chebyshev.chebval((w - pol.cc) / pol.dd, pol.mu)
    ref = Bt @ (U @ (p * (U.T @ z)))
    assert _rel(y, ref) < 1e-10


# ----------------------------------------------------------------- Lanczos
def test_lanbounds_enclose_spectrum(small_p2):
    golden = json.loads((GOLDEN / "tiny.json").read_text())["p2"]
    w = np.array(golden["eigenvalues"])
    P = orc.Pencil(small_p2.A, small_p2.B, bdeg=60)
    lmin, lmax, ns = P.lanbounds(mlan=3000, maxit=5000, tol=1e-5)
    assert lmax >= w[-1] * (1 - 1e-9)
    assert lmax <= w[-1] * 1.01
    assert lmin <= w[0] + 1e-6 * w[-1]
    # bounds of B~ (tol 1e-12) against dense eigenvalues
    Bt = sp.csr_matrix((P.Bt.val, P.Bt.col, P.Bt.rowptr)).toarray()
    eb = np.linalg.eigvalsh(Bt)
    assert abs(P.blmin - eb[0]) < 1e-10 and abs(P.blmax - eb[-1]) < 1e-10


def test_cheblannr_matches_dense_golden(small_p2):
    golden = json.loads((GOLDEN / "tiny.json").read_text())["p2"]
    w = np.array(golden["eigenvalues"])
    lam, X, info = orc.solve_interval(small_p2.A, small_p2.B, (1.0, 12.0), bdeg=60, return_info=True)
    ref = w[(w >= 1.0) & (w <= 12.0)]
    assert lam.size == ref.size == 10
    assert np.max(np.abs(lam - ref) / ref) < 1e-12
    assert np.max(info["res2"]) <= 1e-10
    # eigenvectors are B~-orthonormal and satisfy the pencil
    P = info["pencil"]
    BX = np.stack([P.apply_b(x) for x in X])
    assert np.max(np.abs(X @ BX.T - np.eye(lam.size))) < 1e-10
    # the interval is honoured: [1, 6.5] holds only the lowest multiplet
    lam2, _ = orc.solve_interval(small_p2.A, small_p2.B, (1.0, 6.66), bdeg=60)
    assert lam2.size == np.sum((w >= 1.0) & (w <= 6.66))


def test_cheblannr_reference_degree_residual(small_p2):
    """With the reference B-solve degree (45) the eigenvalues still match to
    1e-12 but the 2-norm residual is limited by the B-solve accuracy."""
    golden = json.loads((GOLDEN / "tiny.json").read_text())["p2"]
    w = np.array(golden["eigenvalues"])
    lam, X, info = orc.solve_interval(small_p2.A, small_p2.B, (1.0, 12.0), bdeg=45, return_info=True)
    ref = w[(w >= 1.0) & (w <= 12.0)]
    assert np.max(np.abs(lam - ref) / ref) < 1e-12
    assert np.max(info["res_ref"]) < 1e-9


def test_p1_golden(small_p1):
    golden = json.loads((GOLDEN / "tiny.json").read_text())["p1"]
    w = np.array(golden["eigenvalues"])
    lam, X, info = orc.solve_interval(small_p1.A, small_p1.B, (1.0, 14.0), bdeg=25, return_info=True)
    ref = w[(w >= 1.0) & (w <= 14.0)]
    assert lam.size == ref.size
    assert np.max(np.abs(lam - ref) / ref) < 1e-12
    assert np.max(info["res2"]) <= 1e-10


def test_start_vector_is_deterministic():
    a = orc.start_vector(1000, 7)
    assert np.array_equal(a, orc.start_vector(1000, 7))
    assert not np.array_equal(a, orc.start_vector(1000, 8))
    assert -1.0 <= a.min() and a.max() < 1.0 and abs(a.mean()) < 0.1


def test_cheblannr_empty_interval_stops():
    """An interval inside a spectral gap holds no eigenvalue: the oracle stops
    once the largest filtered Ritz value has converged below the bar, long
    before maxit, with no pair."""
    import scipy.linalg as sl
    from normalmodes_amd.workloads import build_problem
    P = build_problem("tiny")
    w = sl.eigh(P.A.to_scipy().toarray(), P.B.to_scipy().toarray(), eigvals_only=True)
    nz = w[w > 1.0]
    g = int(np.argmax(np.diff(nz[:30])))
    a, b = nz[g] + 0.3 * (nz[g + 1] - nz[g]), nz[g] + 0.7 * (nz[g + 1] - nz[g])
    lam, X, info = orc.solve_interval(P.A, P.B, (a, b), bdeg=58, return_info=True)
    assert lam.size == 0 and info["nsteps"] < 200


# ------------------------------------------------------------ residual check
def test_residuals_closed_form(small_p2):
    """or_residuals (src/mod_pevsl.f90:137-162) on vectors whose residual is
    known exactly: x = x_j + eps x_k from two B~-orthonormal eigenvectors of
    the scaled pencil gives A~x - lam_j B~x = eps (lam_k - lam_j) B~x_k, so
        ref definition: eps |lam_k - lam_j| ||B~x_k|| / sqrt(N) / |lam_j|
        2-norm gate:    eps |lam_k - lam_j| ||B~x_k|| / sqrt(1 + eps^2) / |lam_j|.
    A dropped /N, a missing B~-normalisation or a residual taken with
    lam B x instead of lam B~x fails one of the two."""
    A, B = small_p2.A.to_scipy(), small_p2.B.to_scipy()
    d = 1.0 / np.sqrt(B.diagonal())
    D = sp.diags(d)
    At = (D @ A @ D).toarray()
    Bt = (D @ B @ D).toarray()
    w, V = sl.eigh(At, Bt)
    j, k = 20, 27
    P = orc.Pencil(small_p2.A, small_p2.B, bdeg=5, bounds_B=(0.2, 5.0))
    n = P.n
    for eps in (0.0, 1e-9, 1e-4, 0.3):
        x = V[:, j] + eps * V[:, k]
        r1, r2 = P.residuals(np.array([w[j]]), x[None, :])
        nb = np.linalg.norm(Bt @ V[:, k])
        want_ref = eps * abs(w[k] - w[j]) * nb / np.sqrt(n) / abs(w[j])
        want_2 = eps * abs(w[k] - w[j]) * nb / np.sqrt(1.0 + eps * eps) / abs(w[j])
        # the dense eigenvector itself leaves a residual of ~1e-13 (rounding)
        floor = 3e-13
        assert abs(r1[0] - want_ref) <= 1e-10 * want_ref + floor, (eps, r1, want_ref)
        assert abs(r2[0] - want_2) <= 1e-10 * want_2 + floor, (eps, r2, want_2)
    # several pairs at once, and a zero eigenvalue is divided by 1, not 0
    X = V[:, [0, j, k]].T.copy()
    r1, r2 = P.residuals(w[[0, j, k]], X)
    assert np.all(r1[1:] < 1e-12) and np.all(r2[1:] < 1e-12)
    r1z, r2z = P.residuals(np.array([0.0]), V[:, [j]].T.copy())
    want = np.linalg.norm(At @ V[:, j])
    assert abs(r2z[0] - want) < 1e-10 * want and abs(r1z[0] - want / np.sqrt(n)) < 1e-10 * want


def test_shift_invert_eigs_closed_form():
    """1D P1 finite elements on n interior nodes, h = 1: K = tridiag(-1, 2, -1),
    M = tridiag(1, 4, 1)/6, eigenvalues 6 (1 - cos t)/(2 + cos t) with
    t = k pi/(n+1).  Near the bottom of this spectrum (lam_max/lam ~ 5e4)
    the shift-invert reference stays within a few 1e-13 relative; LAPACK's
    dense generalized solver, whose absolute error is ~eps ||K||, is several
    times worse; higher up both are near eps."""
    import scipy.sparse as sp
    n = 800
    K = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()
    M = (sp.diags([np.ones(n - 1), 4 * np.ones(n), np.ones(n - 1)], [-1, 0, 1]) / 6.0).tocsr()
    t = np.arange(1, n + 1) * np.pi / (n + 1)
    exact = 12.0 * np.sin(0.5 * t) ** 2 / (2.0 + np.cos(t))    # 1 - cos t without cancellation
    dense = sl.eigh(K.toarray(), M.toarray(), eigvals_only=True)
    for (i0, i1), tol in (((3, 30), 5e-13), ((30, 60), 2e-14)):
        a, b = 0.5 * (exact[i0 - 1] + exact[i0]), 0.5 * (exact[i1] + exact[i1 + 1])
        ref = exact[(exact >= a) & (exact <= b)]
        w = orc.shift_invert_eigs(K, M, a, b)
        assert w.size == ref.size == i1 - i0 + 1
        err = np.max(np.abs(w - ref) / ref)
        assert err < tol, (i0, err)
        wd = orc.shift_invert_eigs(K.toarray(), M.toarray(), a, b)
        assert np.max(np.abs(wd - ref) / ref) < tol
        assert np.max(np.abs(dense[i0:i1 + 1] - ref) / ref) > 3 * err


def test_fluid_pencil_operator_and_filter():
    """The oracle's fluid-solid operator against the dense Schur complement
    of an assembled fluid-core body, and its filter against p applied to the
    dense eigenvalues of the scaled pencil (the Chebyshev solves at degrees
    whose error bound is below 1e-13)."""
    from normalmodes_amd import models
    from normalmodes_amd.assemble import assemble
    from normalmodes_amd.mesh import ball_mesh
    m = models.fluid_core(0.5, radius=6371.0)
    Q = assemble(ball_mesh(600, porder=2, interfaces=m.interfaces, scale=m.radius), m, job=2)
    B = Q.B.to_scipy().toarray()
    Ap = Q.Ap.to_scipy().toarray()
    d = 1.0 / np.sqrt(np.diag(B))
    dp = 1.0 / np.sqrt(-np.diag(Ap))
    Bt = B * d[:, None] * d[None, :]
    Apt = -Ap * dp[:, None] * dp[None, :]
    wb, wa = np.linalg.eigvalsh(Bt), np.linalg.eigvalsh(Apt)

    def degree(lo, hi):
        k = hi / lo
        rho = (np.sqrt(k) - 1) / (np.sqrt(k) + 1)
        return int(np.ceil(np.log(5e-14) / np.log(rho)))
    F = orc.FluidPencil(Q, bdeg=degree(wb[0], wb[-1]), bounds_B=(wb[0], wb[-1]),
                        apdeg=degree(wa[0], wa[-1]), bounds_Ap=(wa[0], wa[-1]))
    S = Q.Ad.to_scipy().toarray() - Q.E.to_scipy().toarray() @ np.linalg.solve(Ap, Q.ET.to_scipy().toarray())
    St = S * d[:, None] * d[None, :]
    x = np.random.default_rng(2).standard_normal(Q.n)
    ref = St @ x
    assert np.linalg.norm(F.apply_a(x) - ref) <= 1e-11 * np.linalg.norm(ref)
    assert np.linalg.norm(Bt @ F.bsolve(x) - x) <= 1e-11 * np.linalg.norm(x)
    # filter: p(A~ B~^-1) acts on B~-eigenvectors' images as p(lambda)
    w, V = sl.eigh(St, Bt)
    lmin, lmax = w[0] - 1e-9 * abs(w[-1]), w[-1] * (1 + 1e-9)
    a, b = w[Q.n // 3], w[Q.n // 3 + 6]
    pol = orc.findpol([a, b, lmin, lmax], 0.8, 0.7)
    z = Bt @ V[:, :40] @ np.ones(40)                       # sum of B~ v_j (images of eigenvectors)
    y = F.chebav(pol, z)
    t = (w[:40] - pol.cc) / pol.dd
    pv_ = np.polynomial.chebyshev.chebval(t, pol.mu)
    ref_y = Bt @ V[:, :40] @ pv_
    assert np.linalg.norm(y - ref_y) <= 1e-8 * np.linalg.norm(z)
