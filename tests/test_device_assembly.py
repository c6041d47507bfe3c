"""The device FE producer (element matrices of CGE3D_ISO on the GPU, gathered
into matrixstruct's CSR pattern) against the CPU producer: same pattern,
values equal to rounding, A exactly symmetric; P1 and P2, JOB=1 and JOB=2."""
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
