"""Per-element eigenvalue sensitivities (the adjoint backward of the path).

d lambda_j / d m_e = x_j^T (dA_e/dm_e - lambda_j dB_e/dm_e) x_j for m_e in
{rho, lambda (Lame), mu} changed uniformly at the nodes of element e, with
x_j the physical, B-normalised eigenvector x = D x~.  The element matrices
are those of the elastic operator (src/mod_cg_create_matrix.f90:1018-1219);
the device computes the quadratic forms directly from the element's nodal
fields (nmg_backward in include/nmgpu.h).  With reference gravity (JOB=2)
the stiffness also depends on rho (OPrho, :1131-1180) with the gravity field
held fixed, and d_rho carries that term (nmg_backward_grav).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import pevsl as pv
from ._native import load_asm, load_gpu
from .refelem import reference_element, vertex_slots

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)

__all__ = ["element_geometry", "element_gravity", "backward", "backward_fs", "pressure_solve", "element_matrices"]


def element_geometry(mesh):
    """detJ (n_elem,) and invJ (n_elem, 3, 3) of the affine element maps."""
    L = load_asm()
    ne, npl = mesh.n_elem, mesh.elems.shape[1]
    detJ = np.empty(ne)
    invJ = np.empty((ne, 9))
    vs = np.asarray(vertex_slots(mesh.porder), dtype=np.int32)
    nodes = np.ascontiguousarray(mesh.nodes)
    elems = np.ascontiguousarray(mesh.elems, dtype=np.int32)
    rc = L.nm_element_geometry(ctypes.c_int64(ne), ctypes.c_int(npl), nodes.ctypes.data_as(_f64p),
                               elems.ctypes.data_as(_i32p), vs.ctypes.data_as(_i32p), detJ.ctypes.data_as(_f64p),
                               invJ.ctypes.data_as(_f64p))
    if rc != 0:
        raise RuntimeError(f"flat element {-rc - 1}")
    return detJ, invJ.reshape(ne, 3, 3)


def element_gravity(mesh, g0):
    """Symmetric least-squares gravity gradients (n_elem, 3, 3) of the nodal
    gravity g0 (n_elem, np, 3), as the stiffness uses them."""
    L = load_asm()
    ne, npl = mesh.n_elem, mesh.elems.shape[1]
    g = np.ascontiguousarray(g0, dtype=np.float64)
    if g.shape != (ne, npl, 3):
        raise ValueError(f"g0 must be (n_elem, np, 3) = {(ne, npl, 3)}, got {g.shape}")
    gsym = np.empty((ne, 9))
    rc = L.nm_element_gravity(ctypes.c_int64(ne), ctypes.c_int(npl),
                              np.ascontiguousarray(mesh.nodes).ctypes.data_as(_f64p),
                              np.ascontiguousarray(mesh.elems, dtype=np.int32).ctypes.data_as(_i32p),
                              g.ctypes.data_as(_f64p), gsym.ctypes.data_as(_f64p))
    if rc != 0:
        raise RuntimeError("element gravity gradients failed")
    return gsym.reshape(ne, 3, 3)


def backward(h, mesh, job: int = 1, g0=None):
    """Sensitivities of the eigenpairs held by solver handle h (after
    apply_pevsl(..., keep=True)).  Returns (d_rho, d_lam, d_mu), each
    (nev, n_elem), rows in ascending eigenvalue order as apply_pevsl returns
    the pairs.  job=2 takes the nodal reference gravity g0 (n_elem, np, 3)
    the stiffness was assembled with."""
    L = load_gpu()
    if job == 2 and g0 is None:
        raise ValueError("job 2 needs g0")
    detJ, invJ = element_geometry(mesh)
    _, M, D = reference_element(mesh.porder)
    ne, npl = mesh.n_elem, mesh.elems.shape[1]
    nev = pv.pevsl_get_nev(h)
    out = [np.empty((nev, ne)) for _ in range(3)]
    elems = np.ascontiguousarray(mesh.elems, dtype=np.int32)
    M = np.ascontiguousarray(M)
    D = np.ascontiguousarray(D)
    invJ = np.ascontiguousarray(invJ.reshape(ne, 9))
    common = (ctypes.c_int64(h), ctypes.c_int64(ne), ctypes.c_int(npl), elems.ctypes.data_as(_i32p),
              detJ.ctypes.data_as(_f64p), invJ.ctypes.data_as(_f64p), M.ctypes.data_as(_f64p),
              D.ctypes.data_as(_f64p))
    outs = tuple(o.ctypes.data_as(_f64p) for o in out)
    if job == 2:
        g = np.ascontiguousarray(g0, dtype=np.float64)
        gsym = np.ascontiguousarray(element_gravity(mesh, g).reshape(ne, 9))
        pv._check(L.nmg_backward_grav(*common, g.ctypes.data_as(_f64p), gsym.ctypes.data_as(_f64p), *outs))
    else:
        pv._check(L.nmg_backward(*common, ctypes.c_int(job), *outs))
    lam, _ = pv.pevsl_copy_result(h, 3 * mesh.n_nodes, False)
    order = np.argsort(lam, kind="stable")
    return tuple(o[order] for o in out)


def element_matrices(mesh, sel, lam, mu, rho, g=None):
    """Element stiffness K (len(sel), 3np, 3np) and scalar mass M (len(sel),
    np, np) from the FE producer for the given per-element-node parameters
    (len(sel), np)."""
    L = load_asm()
    sel = np.ascontiguousarray(sel, dtype=np.int64)
    npl = mesh.elems.shape[1]
    _, Mr, Dr = reference_element(mesh.porder)
    K = np.empty((sel.size, 3 * npl, 3 * npl))
    Me = np.empty((sel.size, npl, npl))
    vs = np.asarray(vertex_slots(mesh.porder), dtype=np.int32)
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (lam, mu, rho)]
    gp = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
    rc = L.nm_element_matrices(ctypes.c_int64(sel.size), ctypes.c_int(npl),
                               np.ascontiguousarray(mesh.nodes).ctypes.data_as(_f64p),
                               np.ascontiguousarray(mesh.elems, dtype=np.int32).ctypes.data_as(_i32p),
                               sel.ctypes.data_as(_i64p), vs.ctypes.data_as(_i32p),
                               *(a.ctypes.data_as(_f64p) for a in args),
                               None if gp is None else gp.ctypes.data_as(_f64p),
                               np.ascontiguousarray(Mr).ctypes.data_as(_f64p),
                               np.ascontiguousarray(Dr).ctypes.data_as(_f64p),
                               K.ctypes.data_as(_f64p), Me.ctypes.data_as(_f64p))
    if rc != 0:
        raise RuntimeError(f"flat element {-rc - 1}")
    return K, Me


def pressure_solve(Ap, R, tol: float = 1e-13, maxit: int = 5000):
    """Ap^-1 R for the negative definite pressure block Ap (CSR) and every
    column of R: conjugate gradients on -Ap with its diagonal as the
    preconditioner, one recurrence per column advanced together.  Raises
    when a column stays above 1e-10 relative residual."""
    A = -Ap
    B = -np.asarray(R, dtype=np.float64).reshape(Ap.shape[0], -1)
    d = A.diagonal()[:, None]
    bn = np.linalg.norm(B, axis=0)
    X = np.zeros_like(B)
    Rr = B.copy()
    Z = Rr / d
    P = Z.copy()
    rz = np.einsum("ij,ij->j", Rr, Z)
    active = bn > 0
    for _ in range(maxit):
        if not active.any():
            break
        AP = A @ P
        pap = np.einsum("ij,ij->j", P, AP)
        alpha = np.where(active, rz / np.where(active, pap, 1.0), 0.0)
        X += P * alpha
        Rr -= AP * alpha
        active &= np.linalg.norm(Rr, axis=0) > tol * bn
        Z = Rr / d
        rz_new = np.einsum("ij,ij->j", Rr, Z)
        beta = np.where(active, rz_new / np.where(rz > 0, rz, 1.0), 0.0)
        P = Z + P * beta
        rz = rz_new
    res = np.linalg.norm(A @ X - B, axis=0) / np.where(bn > 0, bn, 1.0)
    if np.any(res > 1e-10):
        raise RuntimeError(f"pressure solve: relative residual {res.max():.1e} after {maxit} iterations")
    return X


def backward_fs(Q, lam, X, purefluid: bool = False):
    """Sensitivities of eigenpairs of a fluid-solid body (nmg_backward_fs).

    ``Q``: the FluidProblem the pairs belong to; ``lam`` (k,) eigenvalues of
    (Ad - E Ap^-1 ET) u = lambda B u; ``X`` (n, k) the physical displacement
    vectors (B-normalised here).  The pressures p = -Ap^-1 ET u come from
    Jacobi-preconditioned conjugate gradients on -Ap (negative definite) on
    the host, all columns at once.  Returns (d_rho, d_lam, d_mu),
    each (k, n_elem), for rho, lam (Lame) and mu changed uniformly at the
    nodes of each element; d_mu is 0 on fluid elements."""
    from .assemble import face_neighbours
    from .refelem import reference_faces
    L = load_gpu()
    me = Q.mesh
    lam = np.ascontiguousarray(np.atleast_1d(lam), dtype=np.float64)
    X = np.asarray(X, dtype=np.float64).reshape(Q.n, -1)
    k = X.shape[1]
    if lam.size != k:
        raise ValueError("one eigenvalue per vector")
    B = Q.B.to_scipy()
    X = X / np.sqrt(np.einsum("ij,ij->j", X, B @ X))
    Pm = -pressure_solve(Q.Ap.to_scipy(), np.asarray(Q.ET.to_scipy() @ X))
    U = np.ascontiguousarray(X.T)
    P = np.ascontiguousarray(Pm.T)
    npl = me.elems.shape[1]
    _, M, D = reference_element(me.porder)
    fmask, massF = reference_faces(me.porder)
    i8p = ctypes.POINTER(ctypes.c_int8)
    arr = dict(nodes=np.ascontiguousarray(me.nodes, dtype=np.float64),
               elems=np.ascontiguousarray(me.elems, dtype=np.int32),
               vslot=np.asarray(vertex_slots(me.porder), dtype=np.int32),
               lamL=np.ascontiguousarray(Q.lam, dtype=np.float64), rho=np.ascontiguousarray(Q.rho, dtype=np.float64),
               ef=np.ascontiguousarray(Q.fluid.astype(np.int8)), neigh=face_neighbours(me),
               M=np.ascontiguousarray(M), D=np.ascontiguousarray(D), massF=np.ascontiguousarray(massF),
               fmask=np.ascontiguousarray(fmask, dtype=np.int32), vst=np.ascontiguousarray(Q.vstat, dtype=np.int8),
               vstt=np.ascontiguousarray(Q.vstt, dtype=np.int64), pstt=np.ascontiguousarray(Q.pstt, dtype=np.int64))
    g = None if Q.g0 is None else np.ascontiguousarray(Q.g0, dtype=np.float64)
    out = [np.empty((k, me.n_elem)) for _ in range(3)]
    pv._check(L.nmg_backward_fs(
        ctypes.c_int64(me.n_nodes), ctypes.c_int64(me.n_elem), ctypes.c_int(npl), ctypes.c_int(fmask.shape[1]),
        arr["nodes"].ctypes.data_as(_f64p), arr["elems"].ctypes.data_as(_i32p), arr["vslot"].ctypes.data_as(_i32p),
        arr["lamL"].ctypes.data_as(_f64p), arr["rho"].ctypes.data_as(_f64p),
        None if g is None else g.ctypes.data_as(_f64p), arr["ef"].ctypes.data_as(i8p),
        arr["neigh"].ctypes.data_as(_i32p), ctypes.c_int(int(purefluid)), arr["M"].ctypes.data_as(_f64p),
        arr["D"].ctypes.data_as(_f64p), arr["massF"].ctypes.data_as(_f64p), arr["fmask"].ctypes.data_as(_i32p),
        arr["vst"].ctypes.data_as(i8p), arr["vstt"].ctypes.data_as(_i64p), arr["pstt"].ctypes.data_as(_i64p),
        ctypes.c_int(k), lam.ctypes.data_as(_f64p), U.ctypes.data_as(_f64p), P.ctypes.data_as(_f64p),
        *(o.ctypes.data_as(_f64p) for o in out)))
    return tuple(out)
