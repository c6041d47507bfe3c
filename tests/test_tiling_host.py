"""Host-only checks of the LDS row tiling used by the tiled sparse products:
every entry's local index addresses its own column in its tile's node list,
lists fit the capacity and the 16-bit index range, tiles cover all rows."""
import ctypes

import numpy as np
import pytest


def _tiling(nb, rp, ci, cap):
    from normalmodes_amd._native import load_gpu
    L = load_gpu()
    L.nmg_last_error.restype = ctypes.c_char_p
    i64p = ctypes.POINTER(ctypes.c_int64)
    i32p = ctypes.POINTER(ctypes.c_int32)
    u16p = ctypes.POINTER(ctypes.c_uint16)
    nt = ctypes.c_int64()
    mu = ctypes.c_int()
    rc = L.nmg_tiling_host(ctypes.c_int64(nb), rp.ctypes.data_as(i64p), ci.ctypes.data_as(i32p), ctypes.c_int(cap),
                           ctypes.byref(nt), ctypes.byref(mu), None, None, None, None)
    if rc != 0:
        raise ValueError(L.nmg_last_error().decode())
    trp = np.empty(nt.value + 1, np.int64)
    tcp = np.empty(nt.value + 1, np.int64)
    # first pass for the list length
    L.nmg_tiling_host(ctypes.c_int64(nb), rp.ctypes.data_as(i64p), ci.ctypes.data_as(i32p), ctypes.c_int(cap),
                      ctypes.byref(nt), ctypes.byref(mu), trp.ctypes.data_as(i64p), tcp.ctypes.data_as(i64p), None, None)
    tcols = np.empty(tcp[-1], np.int32)
    lcol = np.empty(rp[-1], np.uint16)
    L.nmg_tiling_host(ctypes.c_int64(nb), rp.ctypes.data_as(i64p), ci.ctypes.data_as(i32p), ctypes.c_int(cap),
                      ctypes.byref(nt), ctypes.byref(mu), trp.ctypes.data_as(i64p), tcp.ctypes.data_as(i64p),
                      tcols.ctypes.data_as(i32p), lcol.ctypes.data_as(u16p))
    return trp, tcp, tcols, lcol, mu.value


@pytest.mark.parametrize("cap", [128, 200, 2730])
def test_tiling_invariants(small_p2, cap):
    from normalmodes_amd.assemble import node_adjacency
    rp, ci = node_adjacency(small_p2.mesh)
    nb = rp.size - 1
    trp, tcp, tcols, lcol, mu = _tiling(nb, rp, ci, cap)
    assert trp[0] == 0 and trp[-1] == nb and np.all(np.diff(trp) > 0)
    assert mu <= cap and np.diff(tcp).max() == mu
    for t in range(trp.size - 1):
        lst = tcols[tcp[t]:tcp[t + 1]]
        assert np.all(np.diff(lst) > 0)                       # sorted, distinct
        ks = np.arange(rp[trp[t]], rp[trp[t + 1]])
        assert np.all(lcol[ks] < lst.size)                    # inside the staged list
        assert np.array_equal(lst[lcol[ks]], ci[ks])          # addresses its own column


def test_tiling_rejects_oversized_row(small_p2):
    from normalmodes_amd.assemble import node_adjacency
    rp, ci = node_adjacency(small_p2.mesh)
    with pytest.raises(ValueError):
        _tiling(rp.size - 1, rp, ci, 8)
