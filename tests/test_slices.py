"""Spectrum slices: KPM density of states, the slicer, and the lockstep
multi-slice solve.  The union of the slices' eigenpairs must equal the
single-interval result (no pair lost or duplicated at a boundary)."""
import numpy as np
import pytest


def _exact_tau(t, mdeg):
    return np.array([np.sum(np.cos(k * np.arccos(t))) for k in range(mdeg + 1)])


def test_dos_count_closed_form_vs_quadrature():
    """Closed-form count of the library against the oracle's quadrature of the
    damped density, on random moments (any tau) and on exact moments."""
    from oracle import oracle as orc
    from normalmodes_amd import pevsl as pv
    rng = np.random.default_rng(5)
    for mdeg in (10, 60, 200):
        tau = rng.standard_normal(mdeg + 1) * 10
        tau[0] = 100.0
        for a, b in ((-0.9, 0.3), (-0.2, 0.25), (0.5, 0.99)):
            c = pv.dos_count(tau, [a, b, -1.0, 1.0])
            q = orc.kpm_count(tau, -1.0, 1.0, a, b)
            assert abs(c - q) < 1e-6 * (1 + abs(q)), (mdeg, a, b, c, q)


def test_dos_count_exact_moments_converge():
    """With exact moments of a known spectrum the damped count approaches the
    true count as the degree grows (boundaries placed in gaps)."""
    from normalmodes_amd import pevsl as pv
    rng = np.random.default_rng(11)
    lam = np.sort(rng.uniform(2.0, 30.0, 400))
    lmin, lmax = 1.0, 31.0
    t = (lam - 0.5 * (lmax + lmin)) / (0.5 * (lmax - lmin))
    a = 0.5 * (lam[49] + lam[50])
    b = 0.5 * (lam[149] + lam[150])
    errs = []
    for mdeg in (50, 200, 1000):
        tau = _exact_tau(t, mdeg)
        errs.append(abs(pv.dos_count(tau, [a, b, lmin, lmax]) - 100))
    assert errs[-1] < 3.0 and errs[-1] < errs[0]
    # the full range holds every eigenvalue: count = tau_0 exactly
    tau = _exact_tau(t, 100)
    assert abs(pv.dos_count(tau, [lmin, lmax, lmin, lmax]) - 400) < 1e-9


def test_spslicer_equal_counts():
    from normalmodes_amd import pevsl as pv
    rng = np.random.default_rng(3)
    lam = np.sort(np.concatenate([rng.uniform(2, 8, 300), rng.uniform(8, 30, 100)]))
    lmin, lmax = 1.0, 31.0
    t = (lam - 16.0) / 15.0
    tau = _exact_tau(t, 300)
    intv = [3.0, 25.0, lmin, lmax]
    for ns in (1, 2, 3, 5, 8):
        sli = pv.spslicer(tau, intv, ns)
        assert sli[0] == 3.0 and sli[-1] == 25.0 and np.all(np.diff(sli) > 0)
        total = pv.dos_count(tau, intv)
        for i in range(ns):
            ci = pv.dos_count(tau, [sli[i], sli[i + 1], lmin, lmax])
            assert abs(ci - total / ns) < 1e-6 * total
    # dense region -> narrower slices
    sli = pv.spslicer(tau, intv, 4)
    assert sli[1] - sli[0] < sli[4] - sli[3]


def test_rademacher_block_shape():
    from oracle import oracle as orc
    V = orc.rademacher_block(100, 4, 7, 0)
    assert V.shape == (100, 4) and set(np.unique(V)) == {-1.0, 1.0}
    assert not np.array_equal(V, orc.rademacher_block(100, 4, 7, 1))


# --------------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def mv_small(small_p2):
    from normalmodes_amd.driver import setupmatvec
    mv = setupmatvec(small_p2, bdeg="auto")
    yield mv
    mv.free()


@pytest.mark.gpu
def test_kpm_moments_vs_oracle(small_p2, mv_small):
    """Device moments against the same Rademacher vectors pushed through a
    dense B~ solve; and the count estimate against the true count."""
    from oracle import oracle as orc
    from normalmodes_amd import pevsl as pv
    from normalmodes_amd.driver import _open_solver
    import scipy.linalg as sl
    mv = mv_small
    h = _open_solver(mv, precision=pv.FP64, seed=7, rtol=1e-9)
    try:
        lmin, lmax = pv.pevsl_lanbounds(h, 3000, 5000, 1e-5)
        intv = [1.0, 12.0, lmin, lmax]
        mdeg, nvec = 60, 8
        mu, ecnt = pv.kpmdos(h, mdeg, nvec, intv)
        # the count needs the kernel's resolution at the low end of a wide
        # spectrum (lmax/b ~ 1000): a higher degree
        mu32, _ = pv.kpmdos(h, 400, 32, intv)
    finally:
        pv.pevsl_finish(h)
    d = mv.d
    At = small_p2.A.to_scipy().toarray() * d[:, None] * d[None, :]
    Bt = small_p2.B.to_scipy().toarray() * d[:, None] * d[None, :]
    V = orc.rademacher_block(mv.n, nvec, 7, 0)
    ref = orc.kpm_moments_dense(At, Bt, lmin, lmax, mdeg, V)
    assert np.max(np.abs(mu - ref)) < 1e-9 * mv.n, np.max(np.abs(mu - ref))
    assert abs(mu[0] - mv.n) < 1e-9 * mv.n          # v^T v = n for +-1 vectors
    w = sl.eigh(At, Bt, eigvals_only=True)
    true = int(np.sum((w >= 1.0) & (w <= 12.0)))
    assert abs(pv.dos_count(mu32, intv) - true) < 0.3 * true + 2


@pytest.mark.gpu
@pytest.mark.parametrize("slicer,ns,b", [("uniform", 3, 12.0), ("kpm", 2, 30.0), ("kpm", 4, 60.0),
                                         ("uniform", 1, 12.0), ("kpm", 20, 200.0), ("uniform", 40, 400.0)])
def test_slices_union_equals_single_interval(small_p2, mv_small, dense_eigs, slicer, ns, b):
    """(Narrow slices at the bottom of a wide spectrum need high filter
    degrees, so the KPM cases use wider bands.)"""
    from normalmodes_amd.driver import apply_pevsl, apply_slices
    mv = mv_small
    w = dense_eigs(small_p2)
    ref = w[(w >= 1.0) & (w <= b)]
    r = apply_slices(mv, interval=(1.0, b), n_slices=ns, slicer=slicer, kpm_nvec=16)
    assert r.lam.size == ref.size, (r.lam, ref, r.slices)
    assert np.max(np.abs(r.lam - ref) / ref) < 1e-12
    assert np.all(r.res2 <= 1e-10)
    assert int(np.sum(r.slice_counts)) == r.lam.size
    for i in range(ns):
        inside = np.sum((ref >= r.slices[i]) & (ref <= r.slices[i + 1]))
        assert r.slice_counts[i] == inside
    # the same pairs as the single-interval solve
    s = apply_pevsl(mv, interval=(1.0, b), vectors=False)
    assert np.max(np.abs(s.lam - r.lam) / s.lam) < 1e-12
    # B~-orthonormal union
    B = small_p2.B.to_scipy()
    X = r.X * mv.d[None, :]
    G = X @ (B @ X.T)
    assert np.max(np.abs(G - np.eye(G.shape[0]))) < 1e-8


@pytest.mark.gpu
def test_slice_waves_union_equals_single_interval(small_p2, mv_small, dense_eigs):
    """Eight slices under a basis budget that fits about two at a time: the
    waves run one after another and their union is the single interval's."""
    from normalmodes_amd.driver import apply_pevsl, apply_slices
    mv = mv_small
    w = dense_eigs(small_p2)
    ref = w[(w >= 1.0) & (w <= 60.0)]
    r = apply_slices(mv, interval=(1.0, 60.0), n_slices=8, slicer="uniform", kpm_nvec=16,
                     basis_budget=5.0e6)
    assert len(r.waves) >= 3, r.waves
    assert r.lam.size == ref.size
    assert np.max(np.abs(r.lam - ref) / ref) < 1e-12
    assert np.all(r.res2 <= 1e-10)
    assert int(np.sum(r.slice_counts)) == r.lam.size
    s = apply_pevsl(mv, interval=(1.0, 60.0), vectors=False)
    assert np.max(np.abs(s.lam - r.lam) / s.lam) < 1e-12


@pytest.mark.gpu
def test_slices_boundary_through_multiplet(small_p2, mv_small, dense_eigs):
    """A slice boundary placed exactly on an eigenvalue: the pair is kept once."""
    from normalmodes_amd.driver import apply_slices
    w = dense_eigs(small_p2)
    ref = w[(w >= 1.0) & (w <= 12.0)]
    mid = ref[ref.size // 2]
    r = apply_slices(mv_small, interval=(1.0, 12.0), n_slices=2, slicer="uniform")
    # re-run with the boundary on the eigenvalue
    from normalmodes_amd import pevsl as pv
    from normalmodes_amd.driver import _open_solver
    h = _open_solver(mv_small, precision=pv.FP64, seed=7, rtol=1e-9)
    try:
        lmin, lmax = pv.pevsl_lanbounds(h, 3000, 5000, 1e-5)
        pv.solve_slices(h, [1.0, mid, 12.0], lmin, lmax)
        lam, _ = pv.pevsl_copy_result(h, mv_small.n, False)
    finally:
        pv.pevsl_finish(h)
    lam = np.sort(lam)
    assert lam.size == ref.size and np.max(np.abs(lam - ref) / ref) < 1e-12
    assert r.lam.size == ref.size
