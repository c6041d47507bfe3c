"""The reference's on-disk formats: input mesh and model files, output modes.

Inputs (all little-endian, headerless, in the original global order;
src/mod_geometry.f90:83-425, 1245-1390; src/mod_cg_models.f90:301-429):
  <base>_mesh.header                          ASCII "Ntet Nvert"
  <base>_ele.dat                              int32 [Ntet x 4], 1-based vertex ids
  <base>_neigh.dat                            int32 [Ntet x 4], 1-based neighbour ids, -1 boundary
  <base>_node.dat                             float64 [Nvert x 3], km
  <base>_{vp,vs,rho}_pod_{p}_true.dat         float64 [Ntet x pNp], km/s and g/cm^3,
                                              per element node in the reference's local order
  <base>_pod_{p}_potential_acceleration_true.dat  float64 [Ntet x pNp x 3], m/s^2
Outputs (src/mod_pevsl.f90:188-247; names src/mod_para.f90:271-289):
  <out><base>_JOB<j>_pod<p>_np<n>_<flo>_<fhi>_<i>.dat   float64 [N], mode i
                                              (ascending lambda), physical x = D x~
  <out><base>_pod<p>_np<n>_vlist.dat          int32 [Nnodes], input node id per row node
  <out><base>_pod<p>_np<n>_vstat.dat          int32 [Nnodes], 0 solid (written with fluid only)
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .mesh import TetMesh, mesh_from_vertices

__all__ = ["InputModel", "read_input", "read_mesh", "read_model", "read_gravity", "fortran_real4",
           "mode_file_base", "write_modes", "read_mode"]

_PNP = {1: 4, 2: 10}


@dataclass
class InputModel:
    mesh: TetMesh
    rho: np.ndarray      # (Ntet, pNp) g/cm^3
    vp: np.ndarray       # (Ntet, pNp) km/s
    vs: np.ndarray       # (Ntet, pNp) km/s
    g0: np.ndarray | None  # (Ntet, pNp, 3) m/s^2, if the file is present
    neigh: np.ndarray    # (Ntet, 4) int32, 0-based neighbour ids, -1 boundary


def _path(inputdir, basename, suffix):
    return Path(inputdir) / f"{basename}{suffix}"


def read_mesh(inputdir, basename, porder: int):
    """Vertices and tetrahedra (pnm_read_ele/pnm_read_node) as a P1 or P2
    TetMesh; returns (mesh, neigh)."""
    hdr = _path(inputdir, basename, "_mesh.header").read_text().split()
    ne, nv = int(hdr[0]), int(hdr[1])
    T = np.fromfile(_path(inputdir, basename, "_ele.dat"), dtype="<i4")
    X = np.fromfile(_path(inputdir, basename, "_node.dat"), dtype="<f8")
    if T.size != 4 * ne or X.size != 3 * nv:
        raise ValueError(f"mesh files do not match the header ({ne} tets, {nv} vertices)")
    T = T.reshape(ne, 4).astype(np.int64) - 1
    if T.min() < 0 or T.max() >= nv:
        raise ValueError("vertex id out of range")
    X = X.reshape(nv, 3)
    neigh = None
    p = _path(inputdir, basename, "_neigh.dat")
    if p.exists():
        N = np.fromfile(p, dtype="<i4").reshape(ne, 4)
        neigh = np.where(N > 0, N - 1, -1).astype(np.int32)
        # boundary vertices: those of faces without a neighbour (face j opposite vertex j)
    surf = np.zeros(nv, dtype=bool)
    if neigh is not None:
        for j in range(4):
            f = neigh[:, j] < 0
            for k in range(4):
                if k != j:
                    surf[T[f, k]] = True
    return mesh_from_vertices(X, T, porder, surface=surf), neigh


def read_model(inputdir, basename, porder: int, n_elem: int):
    """rho, vp, vs per element node (src/mod_cg_models.f90:301-360)."""
    out = []
    for q in ("rho", "vp", "vs"):
        a = np.fromfile(_path(inputdir, basename, f"_{q}_pod_{porder}_true.dat"), dtype="<f8")
        if a.size != n_elem * _PNP[porder]:
            raise ValueError(f"{q}: {a.size} values for {n_elem} elements of order {porder}")
        out.append(a.reshape(n_elem, _PNP[porder]))
    return tuple(out)


def read_gravity(inputdir, basename, porder: int, n_elem: int):
    """Reference gravity per element node, m/s^2 (src/mod_cg_models.f90:396-420)."""
    p = _path(inputdir, basename, f"_pod_{porder}_potential_acceleration_true.dat")
    if not p.exists():
        return None
    a = np.fromfile(p, dtype="<f8")
    if a.size != n_elem * _PNP[porder] * 3:
        raise ValueError("gravity file size does not match the mesh")
    return a.reshape(n_elem, _PNP[porder], 3)


def read_input(inputdir, basename, porder: int) -> InputModel:
    mesh, neigh = read_mesh(inputdir, basename, porder)
    rho, vp, vs = read_model(inputdir, basename, porder, mesh.n_elem)
    return InputModel(mesh=mesh, rho=rho, vp=vp, vs=vs, g0=read_gravity(inputdir, basename, porder, mesh.n_elem),
                      neigh=neigh)


def fortran_real4(x: float) -> str:
    """A REAL(4) value as gfortran's list-directed WRITE prints it, left
    adjusted (the reference writes lowfreq/upfreq into file names this way,
    src/mod_para.f90:280-281): 9 significant digits, fixed form for
    0.1 <= |x| < 1e9, else d.dddddddddE+xx."""
    v = float(np.float32(x))
    if v == 0.0:
        return "0.00000000"
    a = abs(v)
    if 0.1 <= a < 1e9:
        digits = 9 - (int(np.floor(np.log10(a))) + 1)
        s = f"{v:.{max(digits, 0)}f}"
        return s
    m, e = f"{v:.8E}".split("E")
    return f"{m}E{int(e):+03d}"


def mode_file_base(outputdir, basename, job: int, porder: int, nproc: int, lowfreq: float, upfreq: float) -> str:
    return (f"{outputdir}{basename}_JOB{job}_pod{porder}_np{nproc}_"
            f"{fortran_real4(lowfreq)}_{fortran_real4(upfreq)}")


def write_modes(outputdir, basename, *, job: int, porder: int, lowfreq: float, upfreq: float, X_phys,
                vlist=None, vstat=None, nproc: int = 1):
    """Mode files in ascending eigenvalue order (row i of X_phys -> _<i+1>.dat,
    float64), and the node list (int32, 1-based input ids) as the reference
    writes them.  Returns the list of paths written."""
    os.makedirs(outputdir, exist_ok=True)
    base = mode_file_base(outputdir, basename, job, porder, nproc, lowfreq, upfreq)
    paths = []
    for i, x in enumerate(np.asarray(X_phys, dtype="<f8")):
        p = f"{base}_{i + 1}.dat"
        x.astype("<f8").tofile(p)
        paths.append(p)
    pre = f"{outputdir}{basename}_pod{porder}_np{nproc}"
    if vlist is not None:
        (np.asarray(vlist, dtype="<i4") + 1).tofile(f"{pre}_vlist.dat")
        paths.append(f"{pre}_vlist.dat")
    if vstat is not None:
        np.asarray(vstat, dtype="<i4").tofile(f"{pre}_vstat.dat")
        paths.append(f"{pre}_vstat.dat")
    return paths


def read_mode(path) -> np.ndarray:
    return np.fromfile(path, dtype="<f8")
