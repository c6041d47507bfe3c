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
from .models import RadialModel, element_node_values, lame
from .refelem import reference_element, vertex_slots

__all__ = ["CSR", "Problem", "assemble", "node_adjacency"]

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


def assemble(mesh: TetMesh, model: RadialModel | None = None, *, rho=None, vp=None, vs=None,
             g0=None, job: int = 1, name: str = "") -> Problem:
    """Assemble A and B.  Give either a model or per-element-node rho, vp, vs.

    ``g0`` (n_elem, pNp, 3) is the reference gravity in km/s^2, used when job == 2.
    """
    if model is not None:
        rho, vp, vs = element_node_values(mesh, model)
    if rho is None:
        raise ValueError("no material")
    if (vs <= 1e-8).all(axis=1).any() if vs.ndim == 2 else False:
        pass
    fluid = vs.sum(axis=1) <= 1e-8 if vs.ndim == 2 else None
    if fluid is not None and fluid.any():
        raise NotImplementedError("fluid elements (vs = 0) need the fluid-solid operator, "
                                  "which this assembler does not build")
    lam, mu = lame(rho, vp, vs)
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
