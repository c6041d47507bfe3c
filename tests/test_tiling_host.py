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


def _cheb2_host(rp, ci, tile_rows):
    import ctypes
    from normalmodes_amd._native import load_gpu
    L = load_gpu()
    nb = rp.size - 1
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    sizes = np.zeros(6, dtype=np.int64)
    P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
    args = [ctypes.c_int64(nb), P(rp, ctypes.c_int64), P(ci, ctypes.c_int32), ctypes.c_int(tile_rows),
            P(sizes, ctypes.c_int64)]
    assert L.nmg_cheb2_host(*args, *([None] * 7)) == 0
    nt, nn, nr, nl = (int(v) for v in sizes[:4])
    out = dict(tile_off=np.zeros(nt + 1, np.int32), row_base=np.zeros(nt + 1, np.int32),
               tile_n=np.zeros(3 * nt, np.int32), nodes=np.zeros(nn, np.int32), row_g=np.zeros(nr, np.int64),
               row_l=np.zeros(nr, np.int32), lcol=np.zeros(max(nl, 1), np.uint16))
    tp = dict(tile_off=ctypes.c_int32, row_base=ctypes.c_int32, tile_n=ctypes.c_int32, nodes=ctypes.c_int32,
              row_g=ctypes.c_int64, row_l=ctypes.c_int32, lcol=ctypes.c_uint16)
    assert L.nmg_cheb2_host(*args, *(P(out[k], tp[k]) for k in tp)) == 0
    out["max_n1"], out["max_n2"] = int(sizes[4]), int(sizes[5])
    return out


@pytest.mark.parametrize("tile_rows", [1, 17, 64])
def test_two_step_tiles_reproduce_two_chebyshev_steps(small_p2, tile_rows):
    """The two-step tiling, run through a host emulation of the tile kernel
    (stage x_k on the 2-hop list, x_{k+1} on the 1-hop list, x_{k+2} on the
    own rows), gives exactly two single Chebyshev steps; every row is owned
    by one tile; local indices stay within the tile lists."""
    B = small_p2.B.to_scipy().tocsr()
    M = B[0::3, 0::3].tocsr()
    M.sort_indices()
    rp, ci, va = M.indptr.astype(np.int64), M.indices.astype(np.int32), M.data
    H = _cheb2_host(rp, ci, tile_rows)
    nb = rp.size - 1
    rng = np.random.default_rng(1)
    xa, xb, b = (rng.standard_normal((nb, 3)) for _ in range(3))
    w1, w2, ith = 1.7, 1.9, 0.37
    xc = np.full((nb, 3), np.nan)
    xd = np.full((nb, 3), np.nan)
    owned = np.zeros(nb, int)
    nt = H["tile_n"].size // 3
    for t in range(nt):
        n0, n1, n2 = H["tile_n"][3 * t:3 * t + 3]
        nd = H["nodes"][H["tile_off"][t]:H["tile_off"][t + 1]]
        assert nd.size == n2 and n0 <= n1 <= n2 and np.unique(nd).size == n2
        rb0 = H["row_base"][t]
        xs = xb[nd]
        xn = np.empty((n1, 3))
        for r in range(n1):
            g0, ls, le = H["row_g"][rb0 + r], H["row_l"][rb0 + r], H["row_l"][rb0 + r + 1]
            cols = H["lcol"][ls:le].astype(int)
            assert cols.max() < n2 and le - ls == rp[nd[r] + 1] - rp[nd[r]] and g0 == rp[nd[r]]
            s = va[g0:g0 + (le - ls)] @ xs[cols]
            u = nd[r]
            xn[r] = xa[u] + w1 * (xs[r] - xa[u] + (b[u] - s) * ith)
        for r in range(n0):
            g0, ls, le = H["row_g"][rb0 + r], H["row_l"][rb0 + r], H["row_l"][rb0 + r + 1]
            cols = H["lcol"][ls:le].astype(int)
            assert cols.max() < n1
            s = va[g0:g0 + (le - ls)] @ xn[cols]
            u = nd[r]
            owned[u] += 1
            xc[u] = xn[r]
            xd[u] = xs[r] + w2 * (xn[r] - xs[r] + (b[u] - s) * ith)
    assert np.all(owned == 1)
    ref_c = xa + w1 * (xb - xa + (b - M @ xb) * ith)
    ref_d = xb + w2 * (ref_c - xb + (b - M @ ref_c) * ith)
    assert np.max(np.abs(xc - ref_c)) < 1e-12 * np.abs(ref_c).max()
    assert np.max(np.abs(xd - ref_d)) < 1e-12 * np.abs(ref_d).max()
