"""The device FE producer against the CPU producer on the same inputs.
Solid bodies: element matrices of CGE3D_ISO on the GPU gathered into
matrixstruct's CSR pattern; same pattern, values equal to rounding, A exactly
symmetric; P1 and P2, JOB=1 and JOB=2.  Fluid-solid bodies: Ad, B, E, ET, Ap
of CGFSE3D_ISO on matrixstruct_general's layout, with interface faces, free
surfaces and the pure-fluid gravity switch."""
import numpy as np
import pytest


def test_pattern_matches_cpu_producer(small_p2, small_p1):
    """The host-side pattern builder used by the device path gives the CPU
    producer's row offsets and columns exactly."""
    from normalmodes_amd.assemble import csr_pattern, node_adjacency
    for P in (small_p2, small_p1):
        ptr, adj = node_adjacency(P.mesh)
        a_rowptr, a_col, b_rowptr, b_col = csr_pattern(ptr, adj)
        assert np.array_equal(a_rowptr, P.A.rowptr) and np.array_equal(a_col, P.A.col)
        assert np.array_equal(b_rowptr, P.B.rowptr) and np.array_equal(b_col, P.B.col)


def _job2_ball():
    from normalmodes_amd.mesh import ball_mesh
    from normalmodes_amd.models import homogeneous
    from normalmodes_amd.assemble import assemble
    me = ball_mesh(600, porder=2, seed=5)
    m = homogeneous()
    g0 = -0.5 * me.nodes[me.elems]            # a uniform sphere's g = -(4/3) pi G rho r, scaled
    return me, m, g0


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["p2", "p1", "layered", "job2"])
def test_device_values_match_cpu(small_p2, small_p1, layered_p2, case):
    from normalmodes_amd.assemble import assemble
    if case == "job2":
        me, m, g0 = _job2_ball()
        ref = assemble(me, m, g0=g0, job=2)
        dev = assemble(me, m, g0=g0, job=2, device=True)
    else:
        ref = {"p2": small_p2, "p1": small_p1, "layered": layered_p2}[case]
        dev = assemble(ref.mesh, rho=ref.rho, vp=np.sqrt((ref.lam + 2 * ref.mu) / ref.rho),
                       vs=np.sqrt(ref.mu / ref.rho), device=True)
        # the CPU producer on the same derived material (vp, vs round trip)
        ref = assemble(ref.mesh, rho=ref.rho, vp=np.sqrt((ref.lam + 2 * ref.mu) / ref.rho),
                       vs=np.sqrt(ref.mu / ref.rho))
    for X, Y in ((dev.A, ref.A), (dev.B, ref.B)):
        assert np.array_equal(X.rowptr, Y.rowptr) and np.array_equal(X.col, Y.col)
        scale = np.abs(Y.val).max()
        err = np.abs(X.val - Y.val).max() / scale
        print(f"{case}: max |device - cpu| / max|val| = {err:.2e}", flush=True)
        assert err < 1e-14
    A = dev.A.to_scipy()
    assert abs(A - A.T).max() == 0.0


# ------------------------------------------------------------ fluid-solid bodies
def _fs_body(case):
    """(mesh, rho, vp, vs, g0, purefluid) of a fluid-solid test body."""
    from normalmodes_amd import models
    from normalmodes_amd.mesh import ball_mesh, stuffed_ball_mesh
    from normalmodes_amd.models import element_node_values, reference_gravity
    if case.startswith("fluidball"):
        # every element fluid: every boundary face is a free surface
        me = ball_mesh(700, porder=2, seed=5)
        sh = me.elems.shape
        rho = 1.5 - 0.3 * np.linalg.norm(me.nodes[me.elems], axis=2)
        return me, rho, 1.7 * np.ones(sh), np.zeros(sh), -0.5 * me.nodes[me.elems], case.endswith("pure")
    m = models.fluid_core(0.5, radius=6371.0) if case.startswith("core") else models.three_layer(6371.0)
    if case == "core-p2-job1":
        me = ball_mesh(900, porder=2, interfaces=m.interfaces, scale=m.radius)
    elif case == "3layer-p1-job2":
        me = stuffed_ball_mesh(20000, porder=1, interfaces=m.interfaces, scale=m.radius)
    else:
        me = stuffed_ball_mesh(6000, porder=2, interfaces=m.interfaces, scale=m.radius)
    rho, vp, vs = element_node_values(me, m)
    g0 = reference_gravity(me, m) if case.endswith("job2") else None
    return me, rho, vp, vs, g0, False


FS_CASES = ["core-p2-job1", "3layer-p2-job2", "3layer-p1-job2", "fluidball-job2", "fluidball-job2-pure"]


@pytest.mark.gpu
@pytest.mark.parametrize("case", FS_CASES)
def test_device_fluid_solid_matches_cpu(case):
    """nmg_assemble_fs against the CPU producer on the same inputs: the same
    columns, values equal to rounding in all five blocks, E = ET^T, and the
    same element statistics (interface elements, fluid and solid elements)."""
    from normalmodes_amd.assemble import assemble_fs
    me, rho, vp, vs, g0, pure = _fs_body(case)
    job = 2 if g0 is not None else 1
    print(f"fluid-solid device producer: {case}, {me.n_elem} tets", flush=True)
    ref = assemble_fs(me, rho=rho, vp=vp, vs=vs, g0=g0, job=job, purefluid=pure)
    dev = assemble_fs(me, rho=rho, vp=vp, vs=vs, g0=g0, job=job, purefluid=pure, device=True)
    assert dev.stats == ref.stats
    scale = {k: np.abs(getattr(ref, k).val).max() for k in ("Ad", "B", "E", "ET", "Ap")}
    if pure:
        # without the N2 terms Ad of an all-fluid body holds only the interior
        # face terms, which cancel between the two elements of each face: it
        # is rounding noise on both sides, measured against the scale of the
        # same body's Ad with its N2 terms
        full = assemble_fs(me, rho=rho, vp=vp, vs=vs, g0=g0, job=job)
        scale["Ad"] = np.abs(full.Ad.val).max()
        assert np.abs(dev.Ad.val).max() <= 1e-12 * scale["Ad"]
    for k in ("Ad", "B", "E", "ET", "Ap"):
        X, Y = getattr(dev, k), getattr(ref, k)
        assert np.array_equal(X.rowptr, Y.rowptr) and np.array_equal(X.col, Y.col), k
        err = np.abs(X.val - Y.val).max() / scale[k]
        print(f"{case} {k}: nnz {Y.nnz}, max |device - cpu| / max|val| = {err:.2e}", flush=True)
        assert err < 1e-13, k
    E, ET = dev.E.to_scipy(), dev.ET.to_scipy()
    assert abs(E - ET.T).max() <= 1e-15 * np.abs(dev.E.val).max()


@pytest.mark.gpu
def test_device_fluid_free_surface_needs_gravity():
    """A fluid element on the free surface without gravity is refused by the
    device producer as by the CPU producer (its surface term divides by the
    mean |g.n|)."""
    from normalmodes_amd.assemble import assemble_fs
    me, rho, vp, vs, _, _ = _fs_body("fluidball-job2")
    with pytest.raises(ValueError, match="free surface"):
        assemble_fs(me, rho=rho, vp=vp, vs=vs, g0=None, job=1)
    with pytest.raises(Exception, match="free surface"):
        assemble_fs(me, rho=rho, vp=vp, vs=vs, g0=None, job=1, device=True)
