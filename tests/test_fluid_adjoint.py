"""Adjoint of fluid-solid bodies (nmg_backward_fs): d lambda / d(rho, lam, mu)
per element for pairs of (Ad - E Ap^-1 ET) u = lambda B u, the pencil of
CGFSE3D_ISO (src/mod_cg_create_matrix.f90:103-977).

Pinned by (a) central finite differences of the eigenvalues of re-assembled
pencils (dense solves of the exact Schur complement) on fluid, interface and
solid elements, JOB=1 and JOB=2; (b) the scaling identities of a body with
element-constant parameters without gravity, where the Schur complement is
homogeneous of degree 1 in (lam, mu) and B of degree 1 in rho:
sum_e (lam_e dl/dlam_e + mu_e dl/dmu_e) = lambda and sum_e rho_e dl/drho_e =
-lambda (P2)."""
import numpy as np
import pytest
import scipy.linalg as sl


def _dense_pairs(Q):
    Ad = Q.Ad.to_scipy().toarray()
    E = Q.E.to_scipy().toarray()
    ET = Q.ET.to_scipy().toarray()
    Ap = Q.Ap.to_scipy().toarray()
    S = Ad - E @ np.linalg.solve(Ap, ET)
    return sl.eigh(0.5 * (S + S.T), Q.B.to_scipy().toarray())


def _simple(w, count):
    """Indices of `count` simple eigenvalues above the fluid's near-zero cluster."""
    out = []
    for i in range(1, len(w) - 1):
        if w[i] > 1e-3 * w[-1] and w[i] - w[i - 1] > 1e-4 * w[i] and w[i + 1] - w[i] > 1e-4 * w[i]:
            out.append(i)
        if len(out) == count:
            break
    return out


def _body(case):
    from normalmodes_amd import models
    from normalmodes_amd.mesh import ball_mesh, stuffed_ball_mesh
    from normalmodes_amd.models import element_node_values, reference_gravity
    name, po, job = case.split("-")
    m = models.fluid_core(0.5, radius=6371.0) if name == "core" else models.three_layer(6371.0)
    if name == "core":
        me = ball_mesh(300 if po == "p1" else 400, porder=int(po[1]), seed=1, interfaces=m.interfaces,
                       scale=m.radius)
    else:
        me = stuffed_ball_mesh(1500, porder=1, interfaces=m.interfaces, scale=m.radius)
    rho, vp, vs = element_node_values(me, m)
    g0 = reference_gravity(me, m) if job == "job2" else None
    return me, rho, vp, vs, g0


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["core-p1-job1", "core-p1-job2", "3layer-p1-job2"])
def test_fluid_adjoint_matches_finite_differences(case):
    from normalmodes_amd.adjoint import backward_fs
    from normalmodes_amd.assemble import assemble_fs
    from normalmodes_amd.models import lame
    me, rho, vp, vs, g0 = _body(case)
    job = 1 if g0 is None else 2
    Q = assemble_fs(me, rho=rho, vp=vp, vs=vs, g0=g0, job=job)
    w, V = _dense_pairs(Q)
    sel = _simple(w, 3)
    assert len(sel) == 3
    print(f"fluid adjoint {case}: n={Q.n}, pairs {sel}", flush=True)
    grads = dict(zip(("rho", "lam", "mu"), backward_fs(Q, w[sel], V[:, sel])))
    assert np.all(grads["mu"][:, Q.fluid] == 0.0)
    lam, mu = lame(rho, vp, vs)
    fl, so = np.nonzero(Q.fluid)[0], np.nonzero(~Q.fluid)[0]
    inter = [e for e in fl if Q.vstat[me.elems[e]].tolist().count(2) >= 3]
    inner = [e for e in fl if e not in inter] or [fl[-1]]
    for e in (inner[0], inter[0], so[0], so[len(so) // 2]):
        for par in ("rho", "lam") + (() if Q.fluid[e] else ("mu",)):
            base = {"rho": rho, "lam": lam, "mu": mu}[par]
            h = 1e-4 * np.abs(base[e]).max()
            vals = []
            for sgn in (1.0, -1.0):
                r2, l2, m2 = rho.copy(), lam.copy(), mu.copy()
                {"rho": r2, "lam": l2, "mu": m2}[par][e] += sgn * h
                Q2 = assemble_fs(me, rho=r2, vp=np.sqrt((l2 + 2 * m2) / r2), vs=np.sqrt(m2 / r2), g0=g0, job=job)
                vals.append(_dense_pairs(Q2)[0][sel])
            fd = (vals[0] - vals[1]) / (2 * h)
            ad = grads[par][:, e]
            err = np.max(np.abs(fd - ad) / np.abs(grads[par]).max(axis=1))
            print(f"  element {e} ({'fluid' if Q.fluid[e] else 'solid'}) d/d{par}: rel err {err:.1e}", flush=True)
            assert err < 1e-4, (e, par, fd, ad)


@pytest.mark.gpu
def test_fluid_adjoint_scaling_identities_p2():
    from normalmodes_amd.adjoint import backward_fs
    from normalmodes_amd.assemble import assemble_fs
    from normalmodes_amd.models import lame
    me, rho, vp, vs, _ = _body("core-p2-job1")
    lam, mu = lame(rho, vp, vs)
    # element-constant parameters: the element means at every node
    rho = np.repeat(rho.mean(axis=1, keepdims=True), rho.shape[1], axis=1)
    lam = np.repeat(lam.mean(axis=1, keepdims=True), lam.shape[1], axis=1)
    mu = np.repeat(mu.mean(axis=1, keepdims=True), mu.shape[1], axis=1)
    Q = assemble_fs(me, rho=rho, vp=np.sqrt((lam + 2 * mu) / rho), vs=np.sqrt(mu / rho))
    w, V = _dense_pairs(Q)
    sel = _simple(w, 4)
    print(f"fluid adjoint scaling identities: n={Q.n}, {int(Q.fluid.sum())} fluid of {me.n_elem} elements",
          flush=True)
    d_rho, d_lam, d_mu = backward_fs(Q, w[sel], V[:, sel])
    stiff = d_lam @ lam[:, 0] + d_mu @ mu[:, 0]
    mass = d_rho @ rho[:, 0]
    e1 = np.max(np.abs(stiff - w[sel]) / w[sel])
    e2 = np.max(np.abs(mass + w[sel]) / w[sel])
    print(f"  sum lam dl/dlam + mu dl/dmu vs lambda: {e1:.1e}; sum rho dl/drho vs -lambda: {e2:.1e}", flush=True)
    assert e1 < 1e-9 and e2 < 1e-9
    # the fluid elements carry part of both sums
    assert np.abs(d_lam[:, Q.fluid]).max() > 1e-6 * np.abs(d_lam).max()


def test_pressure_solve_matches_direct_solve():
    """The host pressure solve p = Ap^-1 r used by the fluid-solid adjoint
    (Jacobi-preconditioned CG, all columns at once) against a sparse direct
    solve, with a zero column among the right-hand sides."""
    import scipy.sparse.linalg as spl
    from normalmodes_amd.adjoint import pressure_solve
    from normalmodes_amd.assemble import assemble_fs
    me, rho, vp, vs, g0 = _body("core-p1-job2")
    Q = assemble_fs(me, rho=rho, vp=vp, vs=vs, g0=g0, job=2)
    Ap = Q.Ap.to_scipy()
    R = np.random.default_rng(4).standard_normal((Q.n_p, 3))
    R[:, 1] = 0.0
    X = pressure_solve(Ap, R)
    ref = spl.spsolve(Ap.tocsc(), R)
    assert np.all(X[:, 1] == 0.0)
    assert np.max(np.abs(X - ref)) <= 1e-11 * np.max(np.abs(ref))
