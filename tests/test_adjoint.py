"""Adjoint sensitivities d lambda / d(rho, lam, mu) per element, on the device,
against (i) x^T (dA_e - lambda dB_e) x formed from the FE producer's own
element matrices on the CPU and (ii) central finite differences of the
eigenvalues of the re-assembled pencil (shift-invert ARPACK)."""
import numpy as np
import pytest

@pytest.fixture(scope="module")
def solved(small_p2):
    from normalmodes_amd.driver import setupmatvec, apply_pevsl
    from normalmodes_amd import pevsl as pv
    from normalmodes_amd.adjoint import backward
    mv = setupmatvec(small_p2, bdeg="auto")
    r = apply_pevsl(mv, interval=(1.0, 12.0), keep=True)
    try:
        g = backward(r.handle, small_p2.mesh)
    finally:
        pv.pevsl_finish(r.handle)
    X = r.X * mv.d[None, :]      # physical, B-normalised
    mv.free()
    return r, g, X


def _reference_gradients(P, X, lam, sel, g0=None):
    """x^T (dK_e - lam dM_e) x from the producer's element matrices: with
    lam = mu = 0 and rho = 1 the element stiffness is exactly the gravity
    part of dK_e/drho_e (linear in rho), zero without gravity."""
    from normalmodes_amd.adjoint import element_matrices
    npl = P.mesh.elems.shape[1]
    one, zero = np.ones((sel.size, npl)), np.zeros((sel.size, npl))
    K_lam, _ = element_matrices(P.mesh, sel, one, zero, zero)
    K_mu, _ = element_matrices(P.mesh, sel, zero, one, zero)
    K_rho, M_rho = element_matrices(P.mesh, sel, zero, zero, one, None if g0 is None else g0[sel])
    out = np.zeros((3, lam.size, sel.size))
    for q, e in enumerate(sel):
        nodes = P.mesh.elems[e]
        dofs = (3 * nodes[None, :] + np.arange(3)[:, None]).ravel()   # (component, node) order of K
        for j in range(lam.size):
            xe = X[j, dofs]
            u = xe.reshape(3, npl)
            out[0, j, q] = xe @ K_rho[q] @ xe - lam[j] * sum(u[p] @ M_rho[q] @ u[p] for p in range(3))
            out[1, j, q] = xe @ K_lam[q] @ xe
            out[2, j, q] = xe @ K_mu[q] @ xe
    return out


GS = 0.5   # gravity of the self-gravitating unit ball: g0 = -GS x


@pytest.fixture(scope="module")
def grav_p2(small_p2):
    """small_p2's mesh and material with reference gravity (JOB=2)."""
    from normalmodes_amd.assemble import assemble
    mesh = small_p2.mesh
    g0 = -GS * mesh.nodes[mesh.elems]                      # (n_elem, np, 3)
    rho, mu, lam = small_p2.rho, small_p2.mu, small_p2.lam
    P = assemble(mesh, rho=rho, vp=np.sqrt((lam + 2 * mu) / rho), vs=np.sqrt(mu / rho), g0=g0, job=2)
    return P, g0


@pytest.mark.gpu
def test_adjoint_matches_element_matrices(small_p2, solved):
    r, g, X = solved
    sel = np.arange(small_p2.mesh.n_elem)
    ref = _reference_gradients(small_p2, X, r.lam, sel)
    for k in range(3):
        scale = np.abs(ref[k]).max()
        assert np.max(np.abs(g[k] - ref[k])) < 1e-11 * scale, k
    # consistency: sum over elements of mu * dlam/dmu + lam * dlam/dlam = lambda
    # (A is linear and homogeneous in (lam, mu) with rho fixed)
    lam_e = small_p2.lam.mean(axis=1)   # homogeneous model: uniform per element
    mu_e = small_p2.mu.mean(axis=1)
    recon = g[1] @ lam_e + g[2] @ mu_e
    assert np.max(np.abs(recon - r.lam) / r.lam) < 1e-9
    # and sum_e rho_e dlam/drho_e = -lambda (B linear in rho)
    rho_e = small_p2.rho.mean(axis=1)
    assert np.max(np.abs(g[0] @ rho_e + r.lam) / r.lam) < 1e-9


@pytest.mark.gpu
def test_adjoint_finite_differences(small_p2, solved, dense_eigs):
    _check_fd(small_p2, solved, dense_eigs)


def test_reference_gradients_finite_differences(small_p2, dense_eigs):
    """The CPU reference (element matrices of the FE producer, quadratic forms
    in the dense eigenvectors) against finite differences, and the Euler
    identities sum_e (lam_e dl/dlam_e + mu_e dl/dmu_e) = l, sum_e rho_e dl/drho_e = -l."""
    import scipy.linalg as sl
    P = small_p2
    w, V = sl.eigh(P.A.to_scipy().toarray(), P.B.to_scipy().toarray())
    sel = np.where((w > 1.0) & (w < 12.0))[0]
    g = _reference_gradients(P, V[:, sel].T, w[sel], np.arange(P.mesh.n_elem))
    lam_e, mu_e, rho_e = P.lam.mean(axis=1), P.mu.mean(axis=1), P.rho.mean(axis=1)
    assert np.max(np.abs(g[1] @ lam_e + g[2] @ mu_e - w[sel]) / w[sel]) < 1e-12
    assert np.max(np.abs(g[0] @ rho_e + w[sel]) / w[sel]) < 1e-12

    class R:
        lam = w[sel]
    _check_fd(P, (R, g, None), dense_eigs)


def test_reference_gradients_gravity_finite_differences(grav_p2):
    """JOB=2: the gravity stiffness is linear in rho with g0 fixed.  The CPU
    reference against central differences of the re-assembled pencil, and
    the Euler identity sum_e (lam_e dl/dlam_e + mu_e dl/dmu_e + rho_e dl/drho_e)
    = x^T A_el x + x^T A_g x - l x^T B x = 0."""
    import scipy.linalg as sl
    P, g0 = grav_p2
    w, V = sl.eigh(P.A.to_scipy().toarray(), P.B.to_scipy().toarray())
    sel = np.where((w > 1.0) & (w < 12.0))[0]
    g = _reference_gradients(P, V[:, sel].T, w[sel], np.arange(P.mesh.n_elem), g0)
    lam_e, mu_e, rho_e = P.lam.mean(axis=1), P.mu.mean(axis=1), P.rho.mean(axis=1)
    assert np.max(np.abs(g[1] @ lam_e + g[2] @ mu_e + g[0] @ rho_e) / w[sel]) < 1e-11
    # the gravity term is not negligible here: without it the density
    # derivatives would sum to -lambda (toroidal modes, with no divergence
    # and little radial motion, feel it least; spheroidal ones most)
    assert np.max(np.abs(g[0] @ rho_e + w[sel]) / w[sel]) > 0.1

    class R:
        lam = w[sel]
    _check_fd(P, (R, g, None), lambda Q: w, g0=g0)


def _check_fd(small_p2, solved, dense_eigs, g0=None):
    import scipy.sparse.linalg as sla
    from normalmodes_amd.assemble import assemble
    r, g, X = solved
    P = small_p2
    # the best-separated computed eigenvalue: its derivative is well defined and
    # the perturbation below moves it far less than its distance to the others
    w = dense_eigs(P)
    gaps = [np.min(np.abs(np.delete(w, np.argmin(np.abs(w - l))) - l)) / l for l in r.lam]
    j = int(np.argmax(gaps))
    lamj = r.lam[j]
    assert gaps[j] > 1e-4
    rho0, lam0, mu0 = P.rho.copy(), P.lam.copy(), P.mu.copy()
    rng = np.random.default_rng(2018)
    top = np.argsort(-np.abs(g[2][j]))[:3]
    elems = np.concatenate([top, rng.choice(P.mesh.n_elem, size=1)])
    rel = 1e-4

    def eig_near(rho, lam, mu):
        vs = np.sqrt(mu / rho)
        vp = np.sqrt((lam + 2 * mu) / rho)
        Q = assemble(P.mesh, rho=rho, vp=vp, vs=vs, g0=g0, job=1 if g0 is None else 2)
        ev = sla.eigsh(Q.A.to_scipy().tocsc(), k=4, M=Q.B.to_scipy().tocsc(), sigma=lamj * (1 - 1e-6),
                       which="LM", tol=1e-14, return_eigenvectors=False)
        return ev[np.argmin(np.abs(ev - lamj))]

    for e in elems:
        for k, base in ((0, rho0), (1, lam0), (2, mu0)):
            h = rel * float(np.abs(base[e]).mean())
            vals = []
            for sgn in (+1, -1):
                arr = [rho0.copy(), lam0.copy(), mu0.copy()]
                arr[k][e] += sgn * h
                vals.append(eig_near(*arr))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - g[k][j, e]) <= 1e-4 * abs(g[k][j, e]) + 1e-7 * lamj, (int(e), k, fd, g[k][j, e])


@pytest.mark.gpu
def test_adjoint_gravity(grav_p2):
    """JOB=2 on the device (nmg_backward_grav) against the producer's element
    matrices, the Euler identity and central differences."""
    from normalmodes_amd.adjoint import backward
    from normalmodes_amd.driver import apply_pevsl, setupmatvec
    from normalmodes_amd import pevsl as pv
    P, g0 = grav_p2
    mv = setupmatvec(P, bdeg="auto")
    try:
        r = apply_pevsl(mv, interval=(1.0, 12.0), keep=True)
        try:
            g = backward(r.handle, P.mesh, job=2, g0=g0)
            with pytest.raises(pv.PevslError, match="nmg_backward_grav"):
                backward(r.handle, P.mesh, job=3, g0=g0)   # nmg_backward takes job 1 only
        finally:
            pv.pevsl_finish(r.handle)
        X = r.X * mv.d[None, :]
    finally:
        mv.free()
    ref = _reference_gradients(P, X, r.lam, np.arange(P.mesh.n_elem), g0)
    for k in range(3):
        scale = np.abs(ref[k]).max()
        assert np.max(np.abs(g[k] - ref[k])) < 1e-11 * scale, k
    lam_e, mu_e, rho_e = P.lam.mean(axis=1), P.mu.mean(axis=1), P.rho.mean(axis=1)
    assert np.max(np.abs(g[1] @ lam_e + g[2] @ mu_e + g[0] @ rho_e) / r.lam) < 1e-9
    import scipy.linalg as sl
    w = sl.eigh(P.A.to_scipy().toarray(), P.B.to_scipy().toarray(), eigvals_only=True)
    _check_fd(P, (r, g, X), lambda Q: w, g0=g0)


@pytest.mark.gpu
def test_adjoint_heterogeneous_ellipsoid():
    """The adjoint on a small version of the heterogeneous-ellipsoid case:
    flattening 1/300, layered profile with a seeded 2% Gaussian random field
    (correlation length 0.1) on rho, vp, vs; the 50 lowest non-rigid modes;
    central finite differences with relative step 1e-6 on 20 seeded random
    elements for the best-separated eigenvalue."""
    import scipy.linalg as sl
    import scipy.sparse.linalg as sla
    from normalmodes_amd.adjoint import backward
    from normalmodes_amd.assemble import assemble
    from normalmodes_amd.driver import apply_pevsl, setupmatvec
    from normalmodes_amd import pevsl as pv
    from normalmodes_amd.mesh import ellipsoid_mesh
    from normalmodes_amd.models import element_node_values, gaussian_random_field, three_layer_solid
    mesh = ellipsoid_mesh(500, flattening=1.0 / 300.0, porder=2, seed=1019)
    rho, vp, vs = element_node_values(mesh, three_layer_solid(), perturb=gaussian_random_field(2018, 0.02, 0.1))
    P = assemble(mesh, rho=rho, vp=vp, vs=vs)
    w = sl.eigh(P.A.to_scipy().toarray(), P.B.to_scipy().toarray(), eigvals_only=True)
    a, b = 0.5 * w[6], 0.5 * (w[55] + w[56])
    ref = w[(w >= a) & (w <= b)]
    assert ref.size == 50
    mv = setupmatvec(P, bdeg="auto")
    try:
        r = apply_pevsl(mv, interval=(a, b), keep=True)
        try:
            g = backward(r.handle, mesh)
        finally:
            pv.pevsl_finish(r.handle)
    finally:
        mv.free()
    assert r.lam.size == 50 and np.max(np.abs(r.lam - ref) / ref) < 1e-10
    lam_n, mu_n, rho_n = P.lam, P.mu, P.rho
    gaps = np.minimum(np.diff(np.concatenate([[0.0], ref])), np.diff(np.concatenate([ref, [np.inf]])))
    j = int(np.argmax(gaps / ref))
    lamj = ref[j]
    rng = np.random.default_rng(2018)
    elems = rng.choice(mesh.n_elem, size=20, replace=False)
    rel = 1e-6

    def eig_near(rh, la, mu):
        v_s = np.sqrt(mu / rh)
        v_p = np.sqrt((la + 2 * mu) / rh)
        Q = assemble(mesh, rho=rh, vp=v_p, vs=v_s)
        ev = sla.eigsh(Q.A.to_scipy().tocsc(), k=4, M=Q.B.to_scipy().tocsc(), sigma=lamj * (1 - 1e-7),
                       which="LM", tol=1e-15, return_eigenvectors=False)
        return ev[np.argmin(np.abs(ev - lamj))]

    for e in elems:
        for k, base in ((0, rho_n), (1, lam_n), (2, mu_n)):
            h = rel * float(np.abs(base[e]).mean())
            vals = []
            for sgn in (+1, -1):
                arr = [rho_n.copy(), lam_n.copy(), mu_n.copy()]
                arr[k][e] += sgn * h
                vals.append(eig_near(*arr))
            fd = (vals[0] - vals[1]) / (2 * h)
            gmax = np.abs(g[k][j]).max()
            assert abs(fd - g[k][j, e]) <= 1e-5 * abs(g[k][j, e]) + 1e-6 * gmax, (int(e), k, fd, g[k][j, e])
