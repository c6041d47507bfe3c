"""Named synthetic workloads (seeded meshes and models).

c1   homogeneous solid unit ball (rho=1, vp=1.73, vs=1), ~12k P2 tets, ~52k DoF
c2s  three-layer radial ball, all layers solid (rho 13->4, vp 11->8,
     vs 3.5->6.5 linear in r), ~466k P2 tets, ~1.9M DoF
c2   the same profiles with a fluid outer core (0.19 < r < 0.55, vs = 0),
     Earth-sized (radius 6371 km, so that self-gravity matters), JOB=2 with
     the model's own reference gravity g0(r) = G M(r)/r^2, interfaces as
     tetrahedron faces, ~466k P2 tets: ~1.97M displacement DoF (six at the
     fluid-solid nodes) and ~100k pressures
c2r  the same body at ~13k tets (~63k DoF), for tests
c4   the same body refined to ~2M tets (~8.6M DoF), for the batched slices
tiny homogeneous ball, ~300 P2 tets (tests and smoke)
c1 and tiny use the jittered-lattice Delaunay mesh; the layered bodies use
the isosurface-stuffed BCC mesh (no slivers: dihedral angles 18-149 degrees,
so the largest eigenvalue stays at the lattice spacing's scale).
Mesh seed 1019, Lanczos start-vector seed 7.
"""
from __future__ import annotations

from .assemble import assemble
from .mesh import ball_mesh, stuffed_ball_mesh
from .models import homogeneous, three_layer, three_layer_solid

MESH_SEED = 1019
LANCZOS_SEED = 7

SPECS = {
    "tiny": dict(n_tets=300, model="homogeneous"),
    "c1": dict(n_tets=12000, model="homogeneous"),
    "c2s": dict(n_tets=450000, model="three_layer_solid"),
    "c2": dict(n_tets=450000, model="three_layer_fluid", job=2),
    "c2r": dict(n_tets=12000, model="three_layer_fluid", job=2),
    "c4": dict(n_tets=2000000, model="three_layer_fluid", job=2),
}


def build_problem(name: str, porder: int = 2, seed: int = MESH_SEED, device: bool = False):
    """Mesh, model and assembled matrices of a workload; ``device`` forms
    the matrices with the GPU producer instead of the CPU one."""
    spec = SPECS[name]
    if spec["model"] == "three_layer_fluid":
        model = three_layer(6371.0)
        mesh = stuffed_ball_mesh(spec["n_tets"], porder=porder, seed=seed, interfaces=model.interfaces,
                                 scale=model.radius)
        return assemble(mesh, model, job=spec.get("job", 1), name=name, device=device)
    if spec["model"] == "three_layer_solid":
        mesh = stuffed_ball_mesh(spec["n_tets"], porder=porder, seed=seed)
        return assemble(mesh, three_layer_solid(), name=name, device=device)
    mesh = ball_mesh(spec["n_tets"], porder=porder, seed=seed)
    return assemble(mesh, homogeneous(), name=name, device=device)

# Eigenvalue intervals (lambda = omega^2 in the model's units).
#   c1: a = lambda_7 / 2 (above the six rigid-body zero modes) and b at the
#   midpoint of the first gap after the 36th eigenvalue, from the shift-invert
#   eigenvalues of the c1 pencil (tests/golden/c1_eigs.json): 33 eigenpairs.
#   c2s: placed by the KPM density of states of the c2s pencil (8 Rademacher
#   vectors, degree 3000) to hold ~200 modes above the six rigid-body modes;
#   the device's FP64 solve finds 231 pairs in it, from the 5-fold 218.42
#   multiplet to 2830.19, with 15.6% and 0.5% margins to the ends.
INTERVALS = {
    "tiny": (1.0, 12.0),
    "c1": (3.1423366861823694, 24.499667699964398),
    "c2s": (189.0, 2845.0),
}
