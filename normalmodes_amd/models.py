"""Material models evaluated per element node.

The reference stores vp, vs and rho per element node (discontinuous across
elements) in km/s and g/cm^3 (src/mod_cg_models.f90:301-360) and takes
lam = rho (vp^2 - 2 vs^2), mu = rho vs^2 (src/mod_cg_create_matrix.f90:1018-1020).
Elements with sum(vs) <= TOL are fluid (separate_fs, src/mod_cg_models.f90:235-277).
This module produces the same per-element-node arrays for the synthetic bodies
used here: a homogeneous ball and radially layered balls whose layer is picked
by the element centroid so that interfaces are sharp, as in the reference data.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .mesh import TetMesh, element_radius

__all__ = ["Layer", "RadialModel", "homogeneous", "three_layer", "three_layer_solid", "fluid_core",
           "element_node_values", "lame", "gaussian_random_field", "mass_within", "reference_gravity",
           "GRAV_CONST"]

# Newton's constant [m^3 kg^-1 s^-2] of the reference's model builder: the
# demo gravity file (CONST_1L_3k.1_pod_1_potential_acceleration_true.dat,
# a homogeneous rho = 5.51 ball) equals (4/3) pi G rho r with this value to
# 3e-12, and differs by >= 4e-5 with any CODATA value.
GRAV_CONST = 6.6723e-11


@dataclass
class Layer:
    r0: float
    r1: float
    rho: tuple[float, float]   # value at r0, value at r1 (linear in r)
    vp: tuple[float, float]
    vs: tuple[float, float]

    def eval(self, r):
        t = (r - self.r0) / (self.r1 - self.r0)
        lin = lambda ab: ab[0] + (ab[1] - ab[0]) * t  # noqa: E731
        return lin(self.rho), lin(self.vp), lin(self.vs)


@dataclass
class RadialModel:
    """Layers over the normalised radius r/radius in [0, 1]; ``radius`` is
    the body's radius in mesh units (km for the reference's models)."""
    layers: list[Layer]
    name: str = "radial"
    axes: tuple[float, float, float] = (1.0, 1.0, 1.0)
    meta: dict = field(default_factory=dict)
    radius: float = 1.0

    @property
    def interfaces(self):
        """Normalised radii where the layers meet."""
        return tuple(L.r1 for L in self.layers[:-1])

    def fluid_layers(self):
        return [k for k, L in enumerate(self.layers) if max(L.vs) == 0.0]

    def layer_index(self, r):
        edges = np.array([L.r1 for L in self.layers[:-1]])
        return np.searchsorted(edges, r, side="right")


def homogeneous(rho=1.0, vp=1.73, vs=1.0) -> RadialModel:
    return RadialModel([Layer(0.0, 1.0, (rho, rho), (vp, vp), (vs, vs))], name="homogeneous")


def _profile(r0, r1, lo, hi):
    # linear in r from the centre value lo (r=0) to the surface value hi (r=1)
    return (lo + (hi - lo) * r0, lo + (hi - lo) * r1)


def three_layer(radius: float = 1.0) -> RadialModel:
    """Solid inner core r<0.19, fluid outer core 0.19-0.55, solid mantle 0.55-1;
    rho 13->4, vp 11->8, vs 3.5->6.5 linear in r (vs = 0 in the fluid).
    ``radius`` in mesh units (6371 for an Earth-sized body in km)."""
    L = []
    for r0, r1, fluid in ((0.0, 0.19, False), (0.19, 0.55, True), (0.55, 1.0, False)):
        rho = _profile(r0, r1, 13.0, 4.0)
        vp = _profile(r0, r1, 11.0, 8.0)
        vs = (0.0, 0.0) if fluid else _profile(r0, r1, 3.5, 6.5)
        L.append(Layer(r0, r1, rho, vp, vs))
    return RadialModel(L, name="three_layer", radius=float(radius))


def fluid_core(r_core: float = 0.5, radius: float = 1.0) -> RadialModel:
    """Two layers: fluid core r < r_core (vs = 0) under a solid mantle, with
    the three-layer profiles of rho, vp and vs.  Meshes of ~1.5k tetrahedra
    resolve both layers (the test body of the fluid-solid operator)."""
    L = []
    for r0, r1, fluid in ((0.0, r_core, True), (r_core, 1.0, False)):
        vs = (0.0, 0.0) if fluid else _profile(r0, r1, 3.5, 6.5)
        L.append(Layer(r0, r1, _profile(r0, r1, 13.0, 4.0), _profile(r0, r1, 11.0, 8.0), vs))
    return RadialModel(L, name="fluid_core", radius=float(radius))


def three_layer_solid() -> RadialModel:
    """The three-layer profile with the outer core kept solid (vs from the same
    linear profile): the layered body on the solid-only operator path."""
    L = []
    for r0, r1 in ((0.0, 0.19), (0.19, 0.55), (0.55, 1.0)):
        L.append(Layer(r0, r1, _profile(r0, r1, 13.0, 4.0), _profile(r0, r1, 11.0, 8.0),
                       _profile(r0, r1, 3.5, 6.5)))
    return RadialModel(L, name="three_layer_solid")


def element_node_values(mesh: TetMesh, model: RadialModel, perturb=None):
    """Return (rho, vp, vs), each (n_elem, pNp) float64.

    ``perturb`` is an optional callable (x[n,3]) -> (drho, dvp, dvs) relative
    perturbations applied multiplicatively at the nodes.
    """
    ax = np.asarray(model.axes, dtype=np.float64) * model.radius
    X = mesh.nodes / ax[None, :]
    xe = X[mesh.elems]                       # (ne, np, 3)
    rn = np.linalg.norm(xe, axis=2)          # node radius
    rc = element_radius(X, np.asarray(mesh.tets, dtype=np.int64))
    li = model.layer_index(rc)
    rho = np.empty_like(rn)
    vp = np.empty_like(rn)
    vs = np.empty_like(rn)
    for k, L in enumerate(model.layers):
        sel = li == k
        if not sel.any():
            continue
        a, b, c = L.eval(rn[sel])
        rho[sel], vp[sel], vs[sel] = a, b, c
    if perturb is not None:
        drho, dvp, dvs = perturb(mesh.nodes[mesh.elems].reshape(-1, 3))
        shp = rho.shape
        rho = rho * (1.0 + drho.reshape(shp))
        vp = vp * (1.0 + dvp.reshape(shp))
        vs = vs * (1.0 + dvs.reshape(shp))
    return rho, vp, vs


def lame(rho, vp, vs):
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2.0 * mu
    return lam, mu


def gaussian_random_field(seed=2018, rms=0.02, corr_len=0.1, n_modes=256):
    """Seeded smooth random field (random Fourier features, Gaussian covariance
    with correlation length ``corr_len``), scaled to ``rms``.  Returns a callable
    x[n,3] -> (drho, dvp, dvs) with independent fields."""
    rng = np.random.default_rng(seed)
    fields = []
    for _ in range(3):
        k = rng.normal(scale=1.0 / corr_len, size=(n_modes, 3))
        ph = rng.uniform(0, 2 * np.pi, size=n_modes)
        fields.append((k, ph))

    def f(x):
        out = []
        for k, ph in fields:
            v = np.sqrt(2.0 / n_modes) * np.cos(x @ k.T + ph).sum(axis=1)
            out.append(rms * v)
        return tuple(out)
    return f


def mass_within(model: RadialModel, r):
    """Mass [kg] inside normalised radius r (array) of a radial model with
    rho in g/cm^3 and ``radius`` in km: 4 pi int_0^r rho(s) s^2 ds, exact for
    the piecewise-linear profiles."""
    r = np.asarray(r, dtype=np.float64)
    Rm = model.radius * 1.0e3
    out = np.zeros_like(r)
    for L in model.layers:
        b = (L.rho[1] - L.rho[0]) / (L.r1 - L.r0)
        a = L.rho[0] - b * L.r0
        lo = L.r0
        hi = np.clip(r, L.r0, L.r1)
        prim = lambda s: a * s ** 3 / 3.0 + b * s ** 4 / 4.0  # noqa: E731
        out += np.where(r > lo, prim(hi) - prim(lo), 0.0)
    return 4.0 * np.pi * out * 1.0e3 * Rm ** 3   # g/cm^3 -> kg/m^3, normalised -> m


def reference_gravity(mesh: TetMesh, model: RadialModel):
    """Reference gravity at every element node, (n_elem, pNp, 3) in km/s^2:
    g0 = -G M(r) r_hat / r^2 of the model's own spherical mass distribution
    (the body's reference state).  The reference reads the same field in
    m/s^2 from ``<base>_pod_<p>_potential_acceleration_true.dat`` and
    divides by 1e3 (src/mod_cg_models.f90:366-429,
    src/mod_cg_create_matrix.f90:194)."""
    if tuple(model.axes) != (1.0, 1.0, 1.0):
        raise ValueError("reference_gravity is for spherical bodies")
    x = mesh.nodes[mesh.elems]                           # (ne, np, 3) mesh units (km)
    r_km = np.linalg.norm(x, axis=2)
    rn = r_km / model.radius
    M = mass_within(model, rn)
    r_m = np.maximum(r_km, 1e-300) * 1.0e3
    gmag = GRAV_CONST * M / r_m ** 2                     # m/s^2
    with np.errstate(invalid="ignore", divide="ignore"):
        rhat = np.where(r_km[..., None] > 0, x / np.maximum(r_km, 1e-300)[..., None], 0.0)
    return -(gmag / 1.0e3)[..., None] * rhat
