"""The hand-written FP64 MFMA kernels of the dense contractions (block
re-orthogonalisation, Ritz vectors, Gram matrices) against numpy, for the
shapes the solver uses and ragged ones (n not a multiple of 16, depth not a
multiple of 4, K from 1 to 64 and beyond), and against the rocBLAS / FMA
path they replace."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pv():
    from normalmodes_amd import pevsl as pv
    return pv


@pytest.mark.parametrize("n,m,K", [(1, 1, 1), (17, 5, 3), (1000, 3, 8), (51540, 424, 8), (20001, 37, 16),
                                   (777, 129, 37), (4099, 64, 64), (3001, 200, 70), (70000, 33, 24)])
def test_gemm_nn_mfma(n, m, K):
    pv = _pv()
    rng = np.random.default_rng(n + m + K)
    A = rng.standard_normal((n, m))
    B = rng.standard_normal((m, K))
    C0 = rng.standard_normal((n, K))
    for alpha, beta in ((1.0, 0.0), (-1.0, 1.0), (0.5, -2.0)):
        C = pv.dense_product(0, 1, A, B, C0, alpha, beta)
        ref = alpha * (A @ B) + beta * C0
        scale = np.abs(A) @ np.abs(B) * abs(alpha) + abs(beta) * np.abs(C0)
        assert np.max(np.abs(C - ref) / (scale + 1e-300)) < 4e-15 * np.sqrt(m), (n, m, K, alpha, beta)


@pytest.mark.parametrize("n,m,K", [(1, 1, 1), (17, 5, 3), (1000, 3, 8), (51540, 424, 8), (20001, 37, 16),
                                   (777, 129, 37), (4099, 300, 64), (3001, 20, 70), (130001, 17, 4)])
def test_gemm_tn_mfma(n, m, K):
    pv = _pv()
    rng = np.random.default_rng(3 * n + m + K)
    A = rng.standard_normal((n, m))
    B = rng.standard_normal((n, K))
    C = pv.dense_product(1, 1, A, B)
    ref = A.T @ B
    scale = np.abs(A).T @ np.abs(B)
    assert np.max(np.abs(C - ref) / scale) < 4e-15 * np.sqrt(n), (n, m, K)
    # the kernels they replace agree
    C0 = pv.dense_product(1, 0, A, B)
    assert np.max(np.abs(C - C0) / scale) < 8e-15 * np.sqrt(n)


def test_block_solve_same_with_either_dense_path(small_p2, dense_eigs):
    """A block-8 solve through the MFMA products finds the same eigenpairs
    as through rocBLAS / the FMA split reduction."""
    from normalmodes_amd.driver import setupmatvec, apply_pevsl
    mv = setupmatvec(small_p2, bdeg="auto")
    try:
        r1 = apply_pevsl(mv, interval=(1.0, 12.0), block=8, dense_kernels=1)
        r0 = apply_pevsl(mv, interval=(1.0, 12.0), block=8, dense_kernels=0)
    finally:
        mv.free()
    w = dense_eigs(small_p2)
    ref = w[(w >= 1.0) & (w <= 12.0)]
    assert r1.lam.size == r0.lam.size == ref.size
    assert np.max(np.abs(r1.lam - r0.lam) / ref) < 1e-13
    assert np.all(r1.res2 <= 1e-10)
