"""The allocation-free Chebyshev solve (nmg_chebiter_noalloc) against the
allocating one (nmg_chebiter) through the C-ABI: two calls in a row on one
device, the second on a smaller matrix, and the size limit of its fixed
scratch array.  Device vectors come from torch (memory plumbing only)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(n_tets, porder):
    from normalmodes_amd.assemble import assemble
    from normalmodes_amd.mesh import ball_mesh
    from normalmodes_amd.models import homogeneous
    return assemble(ball_mesh(n_tets, porder=porder, seed=1019), homogeneous())


def _noalloc(lib, cheb, b_dev, x_dev, stream):
    return lib.nmg_chebiter_noalloc(ctypes.c_int(2), ctypes.c_void_p(b_dev.data_ptr()),
                                    ctypes.c_void_p(x_dev.data_ptr()), ctypes.c_int64(cheb),
                                    ctypes.c_void_p(stream))


def test_noalloc_chebiter_matches_chebiter():
    import torch
    from normalmodes_amd import pevsl as pv
    from normalmodes_amd.driver import setupmatvec
    lib = pv._lib()
    assert torch.cuda.is_available()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev).cuda_stream
    setups = []
    try:
        for n_tets, porder in ((2000, 2), (300, 1)):      # second call: a smaller matrix
            setups.append(setupmatvec(_problem(n_tets, porder), bdeg="auto", tune=()))
        assert setups[1].n < setups[0].n
        torch.cuda.synchronize(dev)
        rng = np.random.default_rng(17)
        for mv in setups:
            b = rng.standard_normal(mv.n)
            ref = pv.pevsl_chebiter(2, b, mv.chebB)
            b_dev = torch.from_numpy(b).to(dev)
            x_dev = torch.full((mv.n,), float("nan"), dtype=torch.float64, device=dev)
            print(f"noalloc chebiter n={mv.n} deg={mv.bdeg}", flush=True)
            assert _noalloc(lib, mv.chebB, b_dev, x_dev, stream) == 0, lib.nmg_last_error()
            torch.cuda.synchronize(dev)
            x = x_dev.cpu().numpy()
            assert np.all(np.isfinite(x))
            assert np.max(np.abs(x - ref)) <= 1e-15 * np.max(np.abs(ref)), np.max(np.abs(x - ref))
            # the same inputs again (the scratch now holds the last call's iterate)
            x_dev.fill_(float("nan"))
            assert _noalloc(lib, mv.chebB, b_dev, x_dev, stream) == 0
            torch.cuda.synchronize(dev)
            assert np.array_equal(x_dev.cpu().numpy(), x)
    finally:
        for mv in setups:
            mv.free()


def test_noalloc_rejects_matrix_beyond_scratch():
    """A matrix with more rows than the scratch array holds is refused before
    any launch."""
    import torch
    from normalmodes_amd import pevsl as pv
    lib = pv._lib()
    n = (1 << 22) + 1
    rowptr = np.arange(n + 1, dtype=np.int64)
    col = np.arange(n, dtype=np.int32)
    val = np.ones(n)
    mat = pv.csr_create(rowptr, col, val)
    cheb = pv.pevsl_setup_chebiter(0.5, 2.0, 3, mat)
    try:
        dev = torch.device("cuda:0")
        b_dev = torch.zeros(4, dtype=torch.float64, device=dev)
        x_dev = torch.zeros(4, dtype=torch.float64, device=dev)
        rc = _noalloc(lib, cheb, b_dev, x_dev, torch.cuda.current_stream(dev).cuda_stream)
        assert rc != 0
    finally:
        pv.pevsl_chebiter_free(cheb)
        pv.csr_free(mat)
