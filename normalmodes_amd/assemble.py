"""Assembly of the generalized eigenproblem A x = lambda B x (solid elastic body).

Python front of ``libnmasm.so`` (nm_assemble.cpp).  Produces CSR matrices in
the layout of the reference's ``CGM%A`` / ``CGM%B`` after setupmatvec's shift
to 0-based columns (src/mod_cg_create_matrix.f90:1269-1455,
src/mod_matvec.f90:47-63): node-major DoFs, int64 row offsets, int32 columns,
float64 values.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._native import load_asm
from .mesh import TetMesh
from .models import RadialModel, element_node_values, lame, reference_gravity
from .refelem import reference_element, reference_faces, vertex_slots

__all__ = ["CSR", "Problem", "FluidProblem", "assemble", "assemble_fs", "node_adjacency", "node_status",
           "face_neighbours"]

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


@dataclass
class CSR:
    n_rows: int
    n_cols: int
    rowptr: np.ndarray   # int64 [n_rows+1], 0-based
    col: np.ndarray      # int32 [nnz], 0-based
    val: np.ndarray      # float64 [nnz]

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.n_rows, self.n_cols))

    def matvec(self, x):
        return self.to_scipy() @ x


@dataclass
class Problem:
    A: CSR
    B: CSR
    mesh: TetMesh
    lam: np.ndarray
    mu: np.ndarray
    rho: np.ndarray
    job: int
    name: str = ""

    @property
    def n(self) -> int:
        return self.A.n_rows


@dataclass
class FluidProblem:
    """A fluid-solid pencil in the reference's blocks (CGM%Ad, %E, %ET, %Ap, %B):
    the operator is Ad - E Ap^-1 ET with Ap negative definite.  Produced by
    ``assemble`` for a body with fluid elements (the mesh and layout fields
    are then set); the blocks alone describe the pencil."""
    Ad: object
    E: object
    ET: object
    Ap: object
    B: object
    porder: int = 2
    name: str = ""
    mesh: TetMesh | None = None
    lam: np.ndarray | None = None
    mu: np.ndarray | None = None
    rho: np.ndarray | None = None
    job: int = 1
    g0: np.ndarray | None = None
    fluid: np.ndarray | None = None    # (n_elem,) bool
    vstat: np.ndarray | None = None    # (n_nodes,) node status 0..5
    vstt: np.ndarray | None = None     # (n_nodes,) first displacement DoF
    pstt: np.ndarray | None = None     # (n_nodes,) pressure DoF or -1
    stats: dict | None = None

    @property
    def n(self) -> int:
        return self.Ad.n_rows

    @property
    def n_p(self) -> int:
        return self.Ap.n_rows


def node_adjacency(mesh: TetMesh):
    lib = load_asm()
    nn, ne, npl = mesh.n_nodes, mesh.n_elem, mesh.elems.shape[1]
    elems = np.ascontiguousarray(mesh.elems, dtype=np.int32)
    ptr = np.zeros(nn + 1, dtype=np.int64)
    lib.nm_node_adjacency(ctypes.c_int64(nn), ctypes.c_int64(ne), ctypes.c_int(npl), _p(elems, _i32p),
                          _p(ptr, _i64p), None)
    adj = np.empty(int(ptr[-1]), dtype=np.int32)
    lib.nm_node_adjacency(ctypes.c_int64(nn), ctypes.c_int64(ne), ctypes.c_int(npl), _p(elems, _i32p),
                          _p(ptr, _i64p), _p(adj, _i32p))
    return ptr, adj


def csr_pattern(ptr, adj):
    """Row offsets and columns of A (3x3 blocks: row 3v+p holds 3w+q for
    every neighbour w of v, ascending) and B (row 3v+p holds 3w+p) from the
    sorted node adjacency: the layout of matrixstruct
    (src/mod_cg_create_matrix.f90:1269-1455)."""
    ptr = np.asarray(ptr, dtype=np.int64)
    adj = np.asarray(adj, dtype=np.int64)
    deg = np.diff(ptr)
    n = deg.size

    def ranges(starts, lens):
        tot = int(lens.sum())
        off = np.concatenate([[0], np.cumsum(lens)[:-1]])
        return np.arange(tot, dtype=np.int64) - np.repeat(off - starts, lens)

    a_len = np.repeat(3 * deg, 3)
    a_rowptr = np.concatenate([[0], np.cumsum(a_len)]).astype(np.int64)
    c3 = (3 * adj[:, None] + np.arange(3)[None, :]).ravel()          # per node, 3 per neighbour
    a_col = c3[ranges(np.repeat(3 * ptr[:-1], 3), a_len)].astype(np.int32)
    b_len = np.repeat(deg, 3)
    b_rowptr = np.concatenate([[0], np.cumsum(b_len)]).astype(np.int64)
    w = adj[ranges(np.repeat(ptr[:-1], 3), b_len)]
    p = np.repeat(np.tile(np.arange(3), n), b_len)
    b_col = (3 * w + p).astype(np.int32)
    return a_rowptr, a_col, b_rowptr, b_col


def assemble(mesh: TetMesh, model: RadialModel | None = None, *, rho=None, vp=None, vs=None,
             g0=None, job: int = 1, name: str = "", device: bool = False) -> Problem:
    """Assemble A and B.  Give either a model or per-element-node rho, vp, vs.

    ``g0`` (n_elem, pNp, 3) is the reference gravity in km/s^2, used when job == 2.
    ``device``: form the matrices on the GPU (``nmg_assemble_elastic`` for a
    solid body, whose pattern is built here; ``nmg_assemble_fs`` for a body
    with fluid elements) instead of the CPU.
    """
    if model is not None:
        rho, vp, vs = element_node_values(mesh, model)
    if rho is None:
        raise ValueError("no material")
    lam, mu = lame(rho, vp, vs)
    fluid = _fluid_elements(vs, mu)
    if fluid.any():
        if job == 2 and g0 is None and model is not None:
            g0 = reference_gravity(mesh, model)
        return assemble_fs(mesh, rho=rho, vp=vp, vs=vs, g0=g0 if job == 2 else None, job=job, name=name,
                           device=device)
    if job == 2 and g0 is None and model is not None and tuple(model.axes) == (1.0, 1.0, 1.0):
        g0 = reference_gravity(mesh, model)
    lib = load_asm()
    npl = mesh.elems.shape[1]
    _, M, D = reference_element(mesh.porder)
    M = np.ascontiguousarray(M)
    D = np.ascontiguousarray(D)
    ptr, adj = node_adjacency(mesh)
    nn, ne = mesh.n_nodes, mesh.n_elem
    nadj = int(ptr[-1])
    a_rowptr = np.empty(3 * nn + 1, dtype=np.int64)
    b_rowptr = np.empty(3 * nn + 1, dtype=np.int64)
    a_col = np.empty(9 * nadj, dtype=np.int32)
    a_val = np.empty(9 * nadj, dtype=np.float64)
    b_col = np.empty(3 * nadj, dtype=np.int32)
    b_val = np.empty(3 * nadj, dtype=np.float64)
    vslot = np.asarray(vertex_slots(mesh.porder), dtype=np.int32)
    g = None
    if job == 2:
        if g0 is None:
            raise ValueError("job 2 needs g0")
        g = np.ascontiguousarray(g0, dtype=np.float64)
    lam_c = np.ascontiguousarray(lam, dtype=np.float64)
    mu_c = np.ascontiguousarray(mu, dtype=np.float64)
    rho_c = np.ascontiguousarray(rho, dtype=np.float64)
    nodes = np.ascontiguousarray(mesh.nodes, dtype=np.float64)
    elems = np.ascontiguousarray(mesh.elems, dtype=np.int32)
    if device:
        from . import pevsl
        a_rowptr, a_col, b_rowptr, b_col = csr_pattern(ptr, adj)
        pevsl.assemble_elastic(nodes, elems, vslot, lam_c, mu_c, rho_c, g, M, D, ptr, adj, a_rowptr, a_val,
                               b_rowptr, b_val)
        n = 3 * nn
        return Problem(A=CSR(n, n, a_rowptr, a_col, a_val), B=CSR(n, n, b_rowptr, b_col, b_val),
                       mesh=mesh, lam=lam, mu=mu, rho=rho, job=job, name=name)
    rc = lib.nm_assemble_elastic(
        ctypes.c_int64(nn), ctypes.c_int64(ne), ctypes.c_int(npl), _p(nodes, _f64p), _p(elems, _i32p),
        _p(vslot, _i32p), _p(lam_c, _f64p), _p(mu_c, _f64p), _p(rho_c, _f64p), _p(g, _f64p),
        _p(M, _f64p), _p(D, _f64p), _p(ptr, _i64p), _p(adj, _i32p),
        _p(a_rowptr, _i64p), _p(a_col, _i32p), _p(a_val, _f64p),
        _p(b_rowptr, _i64p), _p(b_col, _i32p), _p(b_val, _f64p))
    if rc != 0:
        raise RuntimeError(f"assembly failed: flat element {-rc - 1}")
    n = 3 * nn
    return Problem(A=CSR(n, n, a_rowptr, a_col, a_val), B=CSR(n, n, b_rowptr, b_col, b_val),
                   mesh=mesh, lam=lam, mu=mu, rho=rho, job=job, name=name)


def _fluid_elements(vs, mu):
    """Fluid elements.  The reference tests three ways that agree on its
    models: max vs < 1e-6 for the node statuses (src/mod_geometry.f90:286),
    sum vs <= TOL for the element count (src/mod_cg_models.f90:245-247) and
    max mu < TOL in the assembly (src/mod_cg_create_matrix.f90:279), with
    TOL = 1e-8.  A model on which they disagree is rejected."""
    a = vs.max(axis=1) < 1e-6
    b = vs.sum(axis=1) <= 1e-8
    c = mu.max(axis=1) < 1e-8
    if not (np.array_equal(a, b) and np.array_equal(a, c)):
        raise ValueError("the reference's three fluid tests disagree on this model (vs between 0 and 1e-6)")
    return a


def face_neighbours(mesh: TetMesh):
    """neigh[e, f] = element across face f of element e (face f is opposite
    local vertex f), -1 on the boundary: the reference's _neigh.dat
    convention (src/mod_geometry.f90:1273-1278, 2340-2347)."""
    T = np.asarray(mesh.tets, dtype=np.int64)
    ne = T.shape[0]
    faces = np.concatenate([np.sort(np.delete(T, f, axis=1), axis=1) for f in range(4)], axis=0)
    nv = int(T.max()) + 1
    key = (faces[:, 0] * nv + faces[:, 1]) * nv + faces[:, 2]
    order = np.argsort(key, kind="stable")
    ks = key[order]
    neigh = -np.ones(4 * ne, dtype=np.int64)
    same = np.nonzero(ks[1:] == ks[:-1])[0]
    a, b = order[same], order[same + 1]
    neigh[a] = b % ne
    neigh[b] = a % ne
    out = neigh.reshape(4, ne).T
    return np.ascontiguousarray(out.astype(np.int32))


def node_status(mesh: TetMesh, fluid):
    """Node status of the reference: vertices 0 solid, 1 fluid, 2 fluid-solid
    (by the elements that hold them, src/mod_geometry.f90:283-312); P2 edge
    nodes 3, 4, 5 by the smaller status of their two vertices
    (src/mod_geometry.f90:672-680).  That rule is only sound when every edge
    between two fluid-solid vertices touches both a fluid and a solid element,
    every status-3 edge only solid ones and every status-4 edge only fluid
    ones; otherwise the reference's matrices get empty rows, so this raises."""
    nn = mesh.n_nodes
    E = np.asarray(mesh.elems, dtype=np.int64)
    in_f = np.zeros(nn, dtype=bool)
    in_s = np.zeros(nn, dtype=bool)
    in_f[E[fluid].ravel()] = True
    in_s[E[~fluid].ravel()] = True
    st = np.full(nn, -1, dtype=np.int8)
    V = np.asarray(mesh.tets, dtype=np.int64)
    vert = np.unique(V)
    st[vert] = np.where(in_f[vert] & in_s[vert], 2, np.where(in_f[vert], 1, 0))
    if mesh.porder == 2:
        from .mesh import _P2_EDGES
        for a, b, slot in _P2_EDGES:
            node = E[:, slot]
            m = np.minimum(st[V[:, a]], st[V[:, b]])
            st[node] = (m + 3).astype(np.int8)
    if (st < 0).any():
        raise ValueError("node without status")
    six = (st == 2) | (st == 5)
    bad = (six & ~(in_f & in_s)) | ((st == 3) & in_f) | ((st == 4) & in_s)
    if bad.any():
        raise ValueError(f"{int(bad.sum())} nodes whose reference status does not match their elements "
                         "(an edge between two interface vertices inside one region): refine the mesh")
    return st


def assemble_fs(mesh: TetMesh, *, rho, vp, vs, g0=None, job: int = 1, name: str = "",
                purefluid: bool = False, device: bool = False) -> FluidProblem:
    """Fluid-solid assembly (CGFSE3D_ISO, src/mod_cg_create_matrix.f90:103-977,
    on the layout of matrixstruct_general, :1458-2033): Ad, B over the
    displacement DoFs (six at fluid-solid nodes), E, ET and Ap over the
    pressure DoFs.  ``g0`` (n_elem, pNp, 3, km/s^2) switches on the gravity
    terms (JOB=2); without it every gravity term is zero (JOB=1, where the
    reference leaves its local gravity unset).  ``device``: the same
    matrices formed on the GPU (``nmg_assemble_fs``); the layout and row
    offsets are built here either way."""
    lam, mu = lame(rho, vp, vs)
    fluid = _fluid_elements(vs, mu)
    lib = load_asm()
    np_ = mesh.elems.shape[1]
    _, M, D = reference_element(mesh.porder)
    fmask, massF = reference_faces(mesh.porder)
    nfp = fmask.shape[1]
    vst = node_status(mesh, fluid)
    neigh = face_neighbours(mesh)
    ptr, adj = node_adjacency(mesh)
    nn, ne = mesh.n_nodes, mesh.n_elem
    vstt = np.empty(nn, dtype=np.int64)
    pstt = np.empty(nn, dtype=np.int64)
    nd, npr = ctypes.c_int64(), ctypes.c_int64()
    vst = np.ascontiguousarray(vst, dtype=np.int8)
    i8p = ctypes.POINTER(ctypes.c_int8)
    rc = lib.nm_fs_layout(ctypes.c_int64(nn), _p(vst, i8p), _p(vstt, _i64p), _p(pstt, _i64p),
                          ctypes.byref(nd), ctypes.byref(npr))
    if rc != 0:
        raise RuntimeError("bad node status")
    n, n_p = nd.value, npr.value
    rp = {k: np.empty((n if k in ("Ad", "B", "E") else n_p) + 1, dtype=np.int64) for k in ("Ad", "B", "E", "ET", "Ap")}
    rc = lib.nm_fs_rowptr(ctypes.c_int64(nn), _p(ptr, _i64p), _p(adj, _i32p), _p(vst, i8p), _p(vstt, _i64p),
                          _p(pstt, _i64p), *(_p(rp[k], _i64p) for k in ("Ad", "B", "E", "ET", "Ap")))
    if rc != 0:
        raise ValueError(f"fluid-solid pattern: inconsistent node statuses (code {rc})")
    cols = {k: np.empty(int(rp[k][-1]), dtype=np.int32) for k in rp}
    vals = {k: np.empty(int(rp[k][-1]), dtype=np.float64) for k in rp}
    g = None if g0 is None else np.ascontiguousarray(g0, dtype=np.float64)
    ef = np.ascontiguousarray(fluid.astype(np.int8))
    stats = np.zeros(4, dtype=np.int64)
    args = []
    for k in ("Ad", "B", "E", "ET", "Ap"):
        args += [_p(rp[k], _i64p), _p(cols[k], _i32p), _p(vals[k], _f64p)]
    nodes = np.ascontiguousarray(mesh.nodes, dtype=np.float64)
    elems = np.ascontiguousarray(mesh.elems, dtype=np.int32)
    vslot = np.asarray(vertex_slots(mesh.porder), dtype=np.int32)
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (lam, mu, rho, M, D, massF)]
    fmask_c = np.ascontiguousarray(fmask, dtype=np.int32)
    cargs = [ctypes.c_int64(nn), ctypes.c_int64(ne), ctypes.c_int(np_), ctypes.c_int(nfp), _p(nodes, _f64p),
             _p(elems, _i32p), _p(vslot, _i32p), _p(keep[0], _f64p), _p(keep[1], _f64p), _p(keep[2], _f64p),
             _p(g, _f64p), _p(ef, i8p), _p(neigh, _i32p), ctypes.c_int(int(purefluid)),
             _p(keep[3], _f64p), _p(keep[4], _f64p), _p(keep[5], _f64p), _p(fmask_c, _i32p),
             _p(ptr, _i64p), _p(adj, _i32p), _p(vst, i8p), _p(vstt, _i64p), _p(pstt, _i64p), *args,
             _p(stats, _i64p)]
    if device:
        from . import pevsl
        pevsl.assemble_fs(cargs)
        rc = 0
    else:
        rc = lib.nm_assemble_fs(*cargs)
    if rc == -2:
        raise RuntimeError("fluid-solid assembly: an element entry has no slot in the pattern")
    if rc < 0:
        e = -rc - 1
        if e < ne:
            raise RuntimeError(f"assembly failed: flat element {e}")
        raise ValueError(f"fluid element {e - ne} lies on the free surface: the reference's surface term "
                         "divides by the mean |g.n|, which needs gravity (JOB=2)")
    mk = {k: CSR(len(rp[k]) - 1, 0, rp[k], cols[k], vals[k]) for k in rp}
    mk["Ad"].n_cols = mk["B"].n_cols = n
    mk["E"].n_cols = n_p
    mk["ET"].n_cols = n
    mk["Ap"].n_cols = n_p
    return FluidProblem(Ad=mk["Ad"], E=mk["E"], ET=mk["ET"], Ap=mk["Ap"], B=mk["B"], porder=mesh.porder,
                        name=name, mesh=mesh, lam=lam, mu=mu, rho=rho, job=job, g0=g, fluid=fluid, vstat=vst,
                        vstt=vstt, pstt=pstt,
                        stats=dict(interface_elements=int(stats[0]), four_fs_vertices=int(stats[1]),
                                   fluid_elements=int(stats[2]), solid_elements=int(stats[3])))
