"""The allocation-free Chebyshev solve (nmg_chebiter_noalloc) against the
allocating one (nmg_chebiter) through the C-ABI: two calls in a row on one
device, the second on a smaller matrix, and the size limit of its fixed
scratch array.  Device vectors come from the HIP runtime the library
already loaded (hipMalloc / hipMemcpy through ctypes): importing a second
GPU stack into a process that holds this one is not needed and is avoided."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class _Hip:
    """The process's HIP runtime (the one libnmgpu.so is linked to)."""

    def __init__(self):
        from normalmodes_amd import pevsl as pv
        pv._lib()                                   # loads libnmgpu.so and its HIP runtime
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.live = []

    def alloc(self, n):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(8 * n)) == 0
        self.live.append(p)
        return p

    def put(self, dev, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert self.h.hipMemcpy(dev, x.ctypes.data_as(ctypes.c_void_p), x.nbytes, _H2D) == 0

    def get(self, dev, n):
        x = np.empty(n)
        assert self.h.hipMemcpy(x.ctypes.data_as(ctypes.c_void_p), dev, x.nbytes, _D2H) == 0
        return x

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _problem(n_tets, porder):
    from normalmodes_amd.assemble import assemble
    from normalmodes_amd.mesh import ball_mesh
    from normalmodes_amd.models import homogeneous
    return assemble(ball_mesh(n_tets, porder=porder, seed=1019), homogeneous())


def _noalloc(lib, cheb, b_dev, x_dev):
    # the null stream: the call is ordered with the synchronous copies above
    return lib.nmg_chebiter_noalloc(ctypes.c_int(2), b_dev, x_dev, ctypes.c_int64(cheb), ctypes.c_void_p(0))


def test_noalloc_chebiter_matches_chebiter():
    from normalmodes_amd import pevsl as pv
    from normalmodes_amd.driver import setupmatvec
    lib = pv._lib()
    hip = _Hip()
    setups = []
    try:
        for n_tets, porder in ((2000, 2), (300, 1)):      # second call: a smaller matrix
            setups.append(setupmatvec(_problem(n_tets, porder), bdeg="auto", tune=()))
        assert setups[1].n < setups[0].n
        hip.sync()
        rng = np.random.default_rng(17)
        for mv in setups:
            b = rng.standard_normal(mv.n)
            ref = pv.pevsl_chebiter(2, b, mv.chebB)
            b_dev, x_dev = hip.alloc(mv.n), hip.alloc(mv.n)
            hip.put(b_dev, b)
            hip.put(x_dev, np.full(mv.n, np.nan))
            print(f"noalloc chebiter n={mv.n} deg={mv.bdeg}", flush=True)
            assert _noalloc(lib, mv.chebB, b_dev, x_dev) == 0, lib.nmg_last_error()
            hip.sync()
            x = hip.get(x_dev, mv.n)
            assert np.all(np.isfinite(x))
            assert np.max(np.abs(x - ref)) <= 1e-15 * np.max(np.abs(ref)), np.max(np.abs(x - ref))
            # the same inputs again (the scratch now holds the last call's iterate)
            hip.put(x_dev, np.full(mv.n, np.nan))
            assert _noalloc(lib, mv.chebB, b_dev, x_dev) == 0
            hip.sync()
            assert np.array_equal(hip.get(x_dev, mv.n), x)
    finally:
        hip.free()
        for mv in setups:
            mv.free()


def test_noalloc_rejects_matrix_beyond_scratch():
    """A matrix with more rows than the scratch array holds is refused before
    any launch."""
    from normalmodes_amd import pevsl as pv
    lib = pv._lib()
    hip = _Hip()
    n = (1 << 22) + 1
    rowptr = np.arange(n + 1, dtype=np.int64)
    col = np.arange(n, dtype=np.int32)
    val = np.ones(n)
    mat = pv.csr_create(rowptr, col, val)
    cheb = pv.pevsl_setup_chebiter(0.5, 2.0, 3, mat)
    try:
        b_dev, x_dev = hip.alloc(4), hip.alloc(4)
        hip.put(b_dev, np.zeros(4))
        hip.put(x_dev, np.zeros(4))
        rc = _noalloc(lib, cheb, b_dev, x_dev)
        assert rc != 0
    finally:
        hip.free()
        pv.pevsl_chebiter_free(cheb)
        pv.csr_free(mat)
