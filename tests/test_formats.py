"""The reference's file formats: its CONST3k demo input (a data fixture under
tests/golden/const3k) read, meshed and assembled; mode files written and read
back."""
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden" / "const3k"
BASE = "CONST_1L_3k.1"


@pytest.fixture(scope="module")
def const3k():
    from normalmodes_amd.formats import read_input
    return read_input(GOLD, BASE, 1)


def test_read_const3k(const3k):
    M = const3k
    assert M.mesh.n_elem == 3129 and M.mesh.n_nodes == 607
    assert np.all(M.rho == 5.51) and np.all(M.vp == 10.0)
    assert np.allclose(M.vs, 10.0 / np.sqrt(3.0))
    assert M.g0.shape == (3129, 4, 3) and 9.7 < np.abs(M.g0).max() < 9.9
    # every node id of the input appears once; the vertices keep their coordinates
    assert np.array_equal(np.sort(M.mesh.vlist), np.arange(607))
    X = np.fromfile(GOLD / f"{BASE}_node.dat", dtype="<f8").reshape(-1, 3)
    assert np.array_equal(M.mesh.nodes, X[M.mesh.vlist])
    r = np.linalg.norm(M.mesh.nodes, axis=1)
    assert np.allclose(r[M.mesh.surface], 6371.0, rtol=1e-6)
    assert M.neigh.min() == -1 and M.neigh.max() == 3128


def test_const3k_assembly_invariants(const3k):
    from normalmodes_amd.assemble import assemble
    M = const3k
    P = assemble(M.mesh, rho=M.rho, vp=M.vp, vs=M.vs)
    A = P.A.to_scipy()
    B = P.B.to_scipy()
    # total mass per component = rho * polyhedral volume
    T = M.mesh.tets
    X = M.mesh.nodes
    vol = np.abs(np.einsum("ij,ij->i", np.cross(X[T[:, 1]] - X[T[:, 0]], X[T[:, 2]] - X[T[:, 0]]),
                           X[T[:, 3]] - X[T[:, 0]])) / 6.0
    assert abs(B[0::3, 0::3].sum() - 5.51 * vol.sum()) < 1e-12 * 5.51 * vol.sum()
    assert 0.9 < vol.sum() / (4.0 / 3.0 * np.pi * 6371.0 ** 3) < 1.0
    # rigid motions (3 translations, 3 rotations) are in the null space (JOB=1)
    n = P.n
    R = np.zeros((n, 6))
    for c in range(3):
        R[c::3, c] = 1.0
    for k, (i, j) in enumerate(((0, 1), (1, 2), (2, 0))):
        R[i::3, 3 + k] = -X[:, j]
        R[j::3, 3 + k] = X[:, i]
    assert np.abs(A @ R).max() < 1e-9 * np.abs(A).max() * 6371.0
    assert abs(A - A.T).max() < 1e-12 * abs(A).max()


def test_const3k_job2_assembles(const3k):
    from normalmodes_amd.assemble import assemble
    M = const3k
    P = assemble(M.mesh, rho=M.rho, vp=M.vp, vs=M.vs, g0=M.g0 / 1e3, job=2)
    A = P.A.to_scipy()
    assert abs(A - A.T).max() < 1e-10 * abs(A).max()


def test_fortran_real4_names():
    from normalmodes_amd.formats import fortran_real4, mode_file_base
    assert fortran_real4(0.2) == "0.200000003"
    assert fortran_real4(2.0) == "2.00000000"
    assert fortran_real4(10.0) == "10.0000000"
    assert fortran_real4(0.05) == "5.00000007E-02"
    assert mode_file_base("out/", "CONST_1L_3k.1", 1, 1, 1, 0.2, 2.0) == \
        "out/CONST_1L_3k.1_JOB1_pod1_np1_0.200000003_2.00000000"


def test_write_modes_round_trip(tmp_path):
    from normalmodes_amd.formats import read_mode, write_modes
    X = np.random.default_rng(0).standard_normal((3, 30))
    paths = write_modes(str(tmp_path) + "/", "B", job=2, porder=2, lowfreq=0.1, upfreq=1.0, X_phys=X,
                        vlist=np.arange(10)[::-1])
    assert len(paths) == 4
    for i in range(3):
        assert np.array_equal(read_mode(paths[i]), X[i])
        assert paths[i].endswith(f"_{i + 1}.dat")
    assert np.array_equal(np.fromfile(paths[3], dtype="<i4"), np.arange(10)[::-1] + 1)


@pytest.mark.gpu
def test_const3k_demo_solve(tmp_path):
    """The demo configuration (CONST3k, JOB=1, pOrder=1, 0.2-2.0 mHz) through
    the whole path, against the dense pencil; the mode files are B-normalised
    physical vectors."""
    import scipy.linalg as sl
    from normalmodes_amd.driver import freq_to_lambda, solve_input
    from normalmodes_amd.formats import read_mode
    P, mv, r, paths = solve_input(GOLD, BASE, porder=1, job=1, lowfreq=0.2, upfreq=2.0,
                                  outputdir=str(tmp_path) + "/", bdeg="auto")
    a, b = freq_to_lambda(0.2), freq_to_lambda(2.0)
    w = sl.eigh(P.A.to_scipy().toarray(), P.B.to_scipy().toarray(), eigvals_only=True)
    ref = w[(w >= a) & (w <= b)]
    assert ref.size > 5 and r.lam.size == ref.size
    # dense LAPACK carries absolute errors ~eps lam_max (1e-11 relative in
    # this band); the shift-invert eigenvalues of the same pencil pin 1e-12
    from oracle import oracle as orc
    ref_si = orc.shift_invert_eigs(P.A.to_scipy(), P.B.to_scipy(), a, b)
    assert ref_si.size == ref.size
    err = np.max(np.abs(r.lam - ref_si) / ref_si)
    print(f"CONST3k demo: {ref.size} pairs, max rel err vs shift-invert {err:.2e}", flush=True)
    assert err < 1e-12
    assert np.all(r.res2 <= 1e-10)
    B = P.B.to_scipy()
    for i in (0, ref.size - 1):
        x = read_mode(paths[i])
        assert abs(x @ (B @ x) - 1.0) < 1e-9


@pytest.mark.gpu
def test_const3k_job2_adjoint(const3k):
    """The demo model with the reference's own gravity file (JOB=2, P1):
    solve on the device, then d lambda / d(rho, lam, mu) per element through
    nmg_backward_grav against the FE producer's element matrices and the
    Euler identity sum_e (lam_e dl/dlam_e + mu_e dl/dmu_e + rho_e dl/drho_e) = 0
    (elastic stiffness linear in (lam, mu), gravity stiffness and mass linear
    in rho, gravity held fixed)."""
    from normalmodes_amd.adjoint import backward, element_matrices
    from normalmodes_amd.assemble import assemble
    from normalmodes_amd.driver import apply_pevsl, freq_to_lambda, setupmatvec
    from normalmodes_amd import pevsl as pv
    M = const3k
    g0 = M.g0 / 1e3
    P = assemble(M.mesh, rho=M.rho, vp=M.vp, vs=M.vs, g0=g0, job=2)
    mv = setupmatvec(P, bdeg="auto")
    try:
        r = apply_pevsl(mv, interval=(freq_to_lambda(0.2), freq_to_lambda(2.0)), keep=True)
        try:
            g = backward(r.handle, P.mesh, job=2, g0=g0)
        finally:
            pv.pevsl_finish(r.handle)
        X = r.X * mv.d[None, :]
    finally:
        mv.free()
    assert r.lam.size > 0
    ne, npl = P.mesh.n_elem, P.mesh.elems.shape[1]
    sel = np.arange(ne)
    one, zero = np.ones((ne, npl)), np.zeros((ne, npl))
    K_lam, _ = element_matrices(P.mesh, sel, one, zero, zero)
    K_mu, _ = element_matrices(P.mesh, sel, zero, one, zero)
    K_rho, M_rho = element_matrices(P.mesh, sel, zero, zero, one, g0)
    dofs = (3 * P.mesh.elems[:, None, :] + np.arange(3)[None, :, None]).reshape(ne, -1)
    xe = X[:, dofs]                                              # (nev, ne, 3np)
    u = xe.reshape(r.lam.size, ne, 3, npl)
    ref = np.stack([
        np.einsum("jea,eab,jeb->je", xe, K_rho, xe) - r.lam[:, None] * np.einsum("jepa,eab,jepb->je", u, M_rho, u),
        np.einsum("jea,eab,jeb->je", xe, K_lam, xe),
        np.einsum("jea,eab,jeb->je", xe, K_mu, xe)])
    for k in range(3):
        assert np.max(np.abs(g[k] - ref[k])) < 1e-10 * np.abs(ref[k]).max(), k
    s = g[1] @ P.lam.mean(axis=1) + g[2] @ P.mu.mean(axis=1) + g[0] @ P.rho.mean(axis=1)
    assert np.max(np.abs(s) / r.lam) < 1e-9


def test_reference_gravity_matches_demo_file(const3k):
    """The gravity generator on the demo's homogeneous ball (rho 5.51,
    R 6371 km) reproduces the demo's own gravity file, which the reference's
    model builder wrote, at every element node (m/s^2 there, km/s^2 here)."""
    from normalmodes_amd.models import RadialModel, Layer, reference_gravity
    M = const3k
    model = RadialModel([Layer(0.0, 1.0, (5.51, 5.51), (10.0, 10.0), (5.77, 5.77))], radius=6371.0)
    g = reference_gravity(M.mesh, model)
    # surface nodes sit up to 3e-6 km outside R = 6371, where the generator
    # switches to G M / r^2: a 1e-9 relative difference there
    assert np.max(np.abs(g * 1e3 - M.g0)) <= 3e-9 * np.abs(M.g0).max()
    inner = np.linalg.norm(M.mesh.nodes[M.mesh.elems], axis=2) < 6371.0
    assert np.max(np.abs(g * 1e3 - M.g0)[inner]) <= 1e-11 * np.abs(M.g0).max()


def test_face_neighbours_match_demo_file(const3k):
    """Element adjacency across faces (face f opposite local vertex f, -1 on
    the boundary) equals the demo's _neigh.dat."""
    from normalmodes_amd.assemble import face_neighbours
    assert np.array_equal(face_neighbours(const3k.mesh), const3k.neigh)
