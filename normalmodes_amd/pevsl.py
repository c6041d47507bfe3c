"""Python binding of the nmgpu C-ABI, named after the pEVSL Fortran-90 calls
that NormalModes makes (src/mod_pevsl.f90, src/mod_matvec.f90).

Every function raises :class:`PevslError` on a non-zero status; there is no
fallback implementation: the HIP library must be built and a GPU present.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._native import load_gpu

__all__ = ["PevslError", "Handle", "pevsl_start", "pevsl_finish", "pevsl_setprobsizes",
           "pevsl_parcsrcreate", "csr_create", "csr_free", "csr_info", "parcsr_matvec",
           "csr_diagscale", "pevsl_setamv_dad", "pevsl_setamv_fs", "pevsl_setamv", "pevsl_setbmv",
           "pevsl_setbsol_chebiter", "pevsl_set_geneig", "pevsl_set_precision",
           "pevsl_set_seed", "pevsl_setup_chebiter", "pevsl_chebiter", "pevsl_chebiter_free",
           "pevsl_chebiter_stats", "pevsl_lanbounds", "pevsl_findpol", "pol_info",
           "pevsl_freepol", "pevsl_cheblannr", "pevsl_get_nev", "pevsl_copy_result",
           "residuals", "solve_stats", "set_lanczos_controls", "filter_apply", "apply",
           "bench_spmv", "bench_chebstep", "set_kernel_probe", "kernel_probe", "set_block_size", "apply_block", "set_residual_target", "bench_block_chebstep", "set_kernel_map", "set_tiling", "set_xcd_order", "set_stream_loads", "set_value_order", "csr_autotune", "kpmdos", "dos_count", "spslicer",
           "solve_slices", "slice_counts", "set_split_streams", "chebiter_set_mode", "chebiter_info",
           "bench_bsolve", "bench_filter", "chebiter_set_refine", "assemble_elastic", "assemble_fs", "bench_copy", "gemm_tn", "FP64", "FP32_FILTER", "FP32_BSOLVE", "FP32_FILTER_A", "LAYOUTS"]

FP64 = 0
FP32_FILTER = 1
FP32_BSOLVE = 2   # FP32 values and FP32 vectors in the k-column B-solves
FP32_FILTER_A = 3  # FP32 values of B~ in the B-solve and of the stiffness in the filter products
LAYOUTS = {0: "csr", 1: "bsr3", 2: "mass3"}

_i64 = ctypes.c_int64
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_ip = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)
_d = ctypes.c_double


class PevslError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nmgpu error {code}: {msg}")
        self.code = code


_L = None


def _lib():
    global _L
    if _L is None:
        L = load_gpu()
        L.nmg_last_error.restype = ctypes.c_char_p
        _L = L
    return _L


def _check(rc):
    if rc != 0:
        raise PevslError(rc, _lib().nmg_last_error().decode(errors="replace"))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a, t):
    return a.ctypes.data_as(t)


Handle = int


# --- solver handle (src/mod_pevsl.f90:54-57, 220) ----------------------------
def pevsl_start(comm: int = 0) -> Handle:
    h = _i64()
    _check(_lib().nmg_start(ctypes.c_int(comm), ctypes.byref(h)))
    return h.value


def pevsl_finish(h: Handle):
    _check(_lib().nmg_finish(_i64(h)))


def pevsl_setprobsizes(h: Handle, n_glob: int, n_loc: int, nfirst: int = -1):
    _check(_lib().nmg_setprobsizes(_i64(h), _i64(n_glob), _i64(n_loc), _i64(nfirst)))


# --- matrices (src/mod_matvec.f90:69-71 etc.) --------------------------------
def pevsl_parcsrcreate(nrow_g, ncol_g, row_starts, col_starts, ia, ja, a, comm=0) -> Handle:
    ia = np.ascontiguousarray(ia, dtype=np.int32)
    ja = np.ascontiguousarray(ja, dtype=np.int32)
    a = _f64(a)
    rs = np.ascontiguousarray(row_starts, dtype=np.int32)
    cs = np.ascontiguousarray(col_starts, dtype=np.int32)
    m = _i64()
    _check(_lib().nmg_parcsr_create(ctypes.c_int(nrow_g), ctypes.c_int(ncol_g), _p(rs, _ip), _p(cs, _ip),
                                    _p(ia, _ip), _p(ja, _ip), _p(a, _f64p), ctypes.c_int(comm),
                                    ctypes.byref(m)))
    return m.value


def csr_create(rowptr, col, val, n_cols=None) -> Handle:
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = _f64(val)
    n = rowptr.size - 1
    m = _i64()
    _check(_lib().nmg_csr_create(_i64(n), _i64(n if n_cols is None else n_cols), _p(rowptr, _i64p),
                                 _p(col, _i32p), _p(val, _f64p), ctypes.byref(m)))
    return m.value


def csr_free(mat: Handle):
    _check(_lib().nmg_csr_free(_i64(mat)))


def csr_info(mat: Handle):
    nr, nc, nz = _i64(), _i64(), _i64()
    lay = ctypes.c_int()
    _check(_lib().nmg_csr_info(_i64(mat), ctypes.byref(nr), ctypes.byref(nc), ctypes.byref(nz),
                               ctypes.byref(lay)))
    return dict(nrow=nr.value, ncol=nc.value, nnz=nz.value, layout=LAYOUTS[lay.value])


def parcsr_matvec(x, mat: Handle, n_rows=None):
    x = _f64(x)
    info = csr_info(mat) if n_rows is None else dict(nrow=n_rows)
    y = np.empty(info["nrow"])
    _check(_lib().nmg_parcsr_matvec(_p(x, _f64p), _p(y, _f64p), _i64(mat)))
    return y


def csr_diagscale(mat: Handle, sign: float = 1.0):
    n = csr_info(mat)["nrow"]
    d = np.empty(n)
    _check(_lib().nmg_csr_diagscale(_i64(mat), _d(sign), _p(d, _f64p)))
    return d


# --- operators (src/mod_pevsl.f90:69-82) ------------------------------------
def pevsl_setamv_dad(h: Handle, matA: Handle, d):
    d = _f64(d)
    _check(_lib().nmg_set_amv_dad(_i64(h), _i64(matA), _p(d, _f64p)))


def pevsl_setamv_fs(h: Handle, matAd: Handle, matE: Handle, matET: Handle, chebAp: Handle, d, dp):
    """sparsefsAV (src/mod_matvec.f90:498-520): D (Ad - E Ap^-1 ET) D."""
    d = _f64(d)
    dp = _f64(dp)
    _check(_lib().nmg_set_amv_fs(_i64(h), _i64(matAd), _i64(matE), _i64(matET), _i64(chebAp), _p(d, _f64p),
                                 _p(dp, _f64p)))


def pevsl_setamv(h: Handle, mat: Handle):
    _check(_lib().nmg_set_amv(_i64(h), _i64(mat)))


def pevsl_setbmv(h: Handle, matB: Handle):
    _check(_lib().nmg_set_bmv(_i64(h), _i64(matB)))


def pevsl_setbsol_chebiter(h: Handle, chebtype: int, cheb: Handle):
    _check(_lib().nmg_set_bsol_chebiter(_i64(h), ctypes.c_int(chebtype), _i64(cheb)))


def pevsl_set_geneig(h: Handle):
    _check(_lib().nmg_set_geneig(_i64(h)))


def pevsl_set_precision(h: Handle, mode: int):
    _check(_lib().nmg_set_precision(_i64(h), ctypes.c_int(mode)))


def pevsl_set_seed(h: Handle, seed: int):
    _check(_lib().nmg_set_seed(_i64(h), ctypes.c_uint64(seed)))


def set_lanczos_controls(h: Handle, ntest=30, cycle=20, rtol=1e-9):
    _check(_lib().nmg_set_lanczos_controls(_i64(h), ctypes.c_int(ntest), ctypes.c_int(cycle), _d(rtol)))


# --- Chebyshev iteration (src/mod_matvec.f90:88-93, 475-483) ----------------
def pevsl_setup_chebiter(lmin: float, lmax: float, deg: int, mat: Handle) -> Handle:
    c = _i64()
    _check(_lib().nmg_setup_chebiter(_d(lmin), _d(lmax), ctypes.c_int(deg), _i64(mat), ctypes.byref(c)))
    return c.value


def pevsl_chebiter(chebtype: int, b, cheb: Handle):
    b = _f64(b)
    x = np.empty_like(b)
    _check(_lib().nmg_chebiter(ctypes.c_int(chebtype), _p(b, _f64p), _p(x, _f64p), _i64(cheb)))
    return x


def pevsl_chebiter_stats(cheb: Handle):
    a, b = _i64(), _i64()
    s = _d()
    _check(_lib().nmg_chebiter_stats(_i64(cheb), ctypes.byref(a), ctypes.byref(b), ctypes.byref(s)))
    return dict(ncalls=a.value, nmatvec=b.value, seconds=s.value)


def pevsl_chebiter_free(cheb: Handle):
    _check(_lib().nmg_chebiter_free(_i64(cheb)))


# --- eigensolver (src/mod_pevsl.f90:84-130) ---------------------------------
def pevsl_lanbounds(h: Handle, mlan: int, maxit: int, tol: float):
    lo, hi = _d(), _d()
    _check(_lib().nmg_lanbounds(_i64(h), ctypes.c_int(mlan), ctypes.c_int(maxit), _d(tol),
                                ctypes.byref(lo), ctypes.byref(hi)))
    return lo.value, hi.value


def pevsl_findpol(xintv, thr_int: float, thr_ext: float) -> Handle:
    x = _f64(xintv)
    p = _i64()
    _check(_lib().nmg_findpol(_p(x, _f64p), _d(thr_int), _d(thr_ext), ctypes.byref(p)))
    return p.value


@dataclass
class PolInfo:
    deg: int
    cc: float
    dd: float
    gam: float
    bar: float
    mu: np.ndarray


def pol_info(pol: Handle) -> PolInfo:
    deg = ctypes.c_int()
    cc, dd, gam, bar = _d(), _d(), _d(), _d()
    _check(_lib().nmg_pol_info(_i64(pol), ctypes.byref(deg), ctypes.byref(cc), ctypes.byref(dd),
                               ctypes.byref(gam), ctypes.byref(bar), None))
    mu = np.empty(deg.value + 1)
    _check(_lib().nmg_pol_info(_i64(pol), None, None, None, None, None, _p(mu, _f64p)))
    return PolInfo(deg.value, cc.value, dd.value, gam.value, bar.value, mu)


def pevsl_freepol(pol: Handle):
    _check(_lib().nmg_freepol(_i64(pol)))


def pevsl_cheblannr(h: Handle, xintv, maxit: int, tol: float, pol: Handle):
    x = _f64(xintv)
    _check(_lib().nmg_cheblannr(_i64(h), _p(x, _f64p), ctypes.c_int(maxit), _d(tol), _i64(pol)))


def pevsl_get_nev(h: Handle) -> int:
    n = ctypes.c_int()
    _check(_lib().nmg_get_nev(_i64(h), ctypes.byref(n)))
    return n.value


def pevsl_copy_result(h: Handle, n_loc: int, vectors: bool = True):
    nev = pevsl_get_nev(h)
    val = np.empty(nev)
    vec = np.empty((nev, n_loc)) if vectors else None
    _check(_lib().nmg_copy_result(_i64(h), _p(val, _f64p), _p(vec, _f64p) if vectors else None, _i64(n_loc)))
    return val, vec


def residuals(h: Handle):
    nev = pevsl_get_nev(h)
    r1 = np.empty(nev)
    r2 = np.empty(nev)
    _check(_lib().nmg_residuals(_i64(h), _p(r1, _f64p), _p(r2, _f64p)))
    return r1, r2


def solve_stats(h: Handle):
    ls = ctypes.c_int()
    fa = _i64()
    nc = ctypes.c_int()
    sec = _d()
    _check(_lib().nmg_solve_stats(_i64(h), ctypes.byref(ls), ctypes.byref(fa), ctypes.byref(nc),
                                  ctypes.byref(sec)))
    return dict(lan_steps=ls.value, filter_applies=fa.value, ncand=nc.value, seconds=sec.value)


def solve_quality(h: Handle):
    """(target_met, resid_max) of the residual stop of the last solve."""
    tm = ctypes.c_int()
    rm = _d()
    _check(_lib().nmg_solve_quality(_i64(h), ctypes.byref(tm), ctypes.byref(rm)))
    return bool(tm.value), rm.value


def kpmdos(h: Handle, mdeg: int, nvec: int, intv):
    """Trace moments mu[0..mdeg] of T_k((A~B~^-1 - cc)/dd) and the estimated
    eigenvalue count in [intv[0], intv[1]] (pEVSL kpmdos)."""
    x = _f64(intv)
    mu = np.empty(mdeg + 1)
    ecnt = _d()
    _check(_lib().nmg_kpmdos(_i64(h), ctypes.c_int(mdeg), ctypes.c_int(nvec), _p(x, _f64p), _p(mu, _f64p),
                             ctypes.byref(ecnt)))
    return mu, ecnt.value


def dos_count(mu, intv) -> float:
    m = _f64(mu)
    x = _f64(intv)
    c = _d()
    _check(_lib().nmg_dos_count(_p(m, _f64p), ctypes.c_int(m.size - 1), _p(x, _f64p), ctypes.byref(c)))
    return c.value


def spslicer(mu, intv, n_slices: int) -> np.ndarray:
    """Slice boundaries with equal estimated counts (pEVSL spslicer)."""
    m = _f64(mu)
    x = _f64(intv)
    sli = np.empty(n_slices + 1)
    _check(_lib().nmg_spslicer(_p(sli, _f64p), _p(m, _f64p), ctypes.c_int(m.size - 1), _p(x, _f64p),
                               ctypes.c_int(n_slices)))
    return sli


def solve_slices(h: Handle, sli, lmin: float, lmax: float, thr_int: float = 0.8, thr_ext: float = 0.7,
                 maxit: int = 9624, tol: float = 1e-5):
    s = _f64(sli)
    _check(_lib().nmg_solve_slices(_i64(h), ctypes.c_int(s.size - 1), _p(s, _f64p), _d(lmin), _d(lmax),
                                   _d(thr_int), _d(thr_ext), ctypes.c_int(maxit), _d(tol)))


def slice_counts(h: Handle, n_slices: int) -> np.ndarray:
    c = np.zeros(n_slices, dtype=np.int32)
    _check(_lib().nmg_slice_counts(_i64(h), _p(c, _i32p)))
    return c


def set_verbose(h: Handle, every: int):
    """Progress line on stderr every `every` Lanczos steps (0: silent)."""
    _check(_lib().nmg_set_verbose(_i64(h), ctypes.c_int(int(every))))


def set_split_filter(h: Handle, on: bool = True):
    """With split streams, run whole filter chains per part (default) or
    split only the B-solves."""
    _check(_lib().nmg_set_split_filter(_i64(h), ctypes.c_int(int(bool(on)))))


def set_graphs(h: Handle, on: bool = True):
    """Replay B-solves from captured HIP graphs (one host launch per solve)."""
    _check(_lib().nmg_set_graphs(_i64(h), ctypes.c_int(int(bool(on)))))


def set_split_streams(h: Handle, parts: int = 2):
    """k-column B-solves in `parts` (2 or 4) concurrent parts on as many streams; 0 turns it off."""
    _check(_lib().nmg_set_split_streams(_i64(h), ctypes.c_int(int(parts))))


def set_residual_target(h: Handle, target: float):
    _check(_lib().nmg_set_residual_target(_i64(h), _d(target)))


def set_block_size(h: Handle, block: int):
    _check(_lib().nmg_set_block_size(_i64(h), ctypes.c_int(block)))


def set_kernel_probe(h: Handle, every: int):
    _check(_lib().nmg_set_kernel_probe(_i64(h), ctypes.c_int(every)))


def kernel_probe(h: Handle, which: int):
    """which: 0 fused Chebyshev B-solve step, 1 fused filter step.
    Returns (mean seconds per launch, samples, algorithmic bytes per launch)."""
    s, b = _d(), _d()
    n = _i64()
    _check(_lib().nmg_kernel_probe(_i64(h), ctypes.c_int(which), ctypes.byref(s), ctypes.byref(n),
                                   ctypes.byref(b)))
    return s.value, n.value, b.value


def filter_apply(h: Handle, pol: Handle, z):
    z = _f64(z)
    y = np.empty_like(z)
    _check(_lib().nmg_filter_apply(_i64(h), _i64(pol), _p(z, _f64p), _p(y, _f64p)))
    return y


def apply(h: Handle, which: int, x):
    """which: 0 A~ (DAD), 1 B~, 2 B-solve."""
    x = _f64(x)
    y = np.empty_like(x)
    _check(_lib().nmg_apply(_i64(h), ctypes.c_int(which), _p(x, _f64p), _p(y, _f64p)))
    return y


def assemble_elastic(nodes, elems, vslot, lam, mu, rho, g, refM, refD, adj_ptr, adj, a_rowptr, a_val,
                     b_rowptr, b_val):
    """Solid-region A and B values on the GPU (nmg_assemble_elastic); a_val
    and b_val (preallocated, float64) are filled in place."""
    n_nodes = adj_ptr.size - 1
    n_elem, np_ = elems.shape
    for a, t in ((a_val, np.float64), (b_val, np.float64)):
        if a.dtype != t or not a.flags.c_contiguous:
            raise ValueError("a_val / b_val must be contiguous float64")
    _check(_lib().nmg_assemble_elastic(
        _i64(n_nodes), _i64(n_elem), ctypes.c_int(np_), _p(nodes, _f64p), _p(elems, _i32p), _p(vslot, _i32p),
        _p(lam, _f64p), _p(mu, _f64p), _p(rho, _f64p), None if g is None else _p(g, _f64p), _p(refM, _f64p),
        _p(refD, _f64p), _p(adj_ptr, _i64p), _p(adj, _i32p), _p(a_rowptr, _i64p), _p(a_val, _f64p),
        _p(b_rowptr, _i64p), _p(b_val, _f64p)))


def assemble_fs(args):
    """Fluid-solid Ad, B, E, ET, Ap on the GPU (nmg_assemble_fs): ``args`` is
    the ctypes argument list of the CPU producer nm_assemble_fs, whose
    column and value arrays are filled in place."""
    _check(_lib().nmg_assemble_fs(*args))


def chebiter_set_refine(cheb: Handle, passes: int, inner_deg: int = 0):
    """Mixed-precision refinement of a Chebyshev solve: ``passes`` inner FP32
    solves of degree ``inner_deg`` on FP64 residuals (0 passes: plain FP64)."""
    _check(_lib().nmg_chebiter_set_refine(_i64(cheb), ctypes.c_int(passes), ctypes.c_int(inner_deg)))


def bench_filter(h: Handle, pol: Handle, K: int = 1, reps: int = 3) -> float:
    """Mean seconds of one filter application to K device-resident columns."""
    sec = _d()
    _check(_lib().nmg_bench_filter(_i64(h), _i64(pol), ctypes.c_int(K), ctypes.c_int(reps), ctypes.byref(sec)))
    return sec.value


def apply_block(h: Handle, which: int, X):
    """X: (K, n) host block; returns op applied to every row (column of the block)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.empty_like(X)
    _check(_lib().nmg_apply_block(_i64(h), ctypes.c_int(which), ctypes.c_int(X.shape[0]), _p(X, _f64p),
                                  _p(Y, _f64p)))
    return Y


def chebiter_set_mode(cheb: Handle, mode: int):
    """0 (default): one launch per step; 1: one persistent launch per solve where it applies."""
    _check(_lib().nmg_chebiter_set_mode(_i64(cheb), ctypes.c_int(mode)))


def chebiter_info(cheb: Handle, fp32: bool = False):
    p, g, l = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(_lib().nmg_chebiter_info(_i64(cheb), ctypes.c_int(int(fp32)), ctypes.byref(p), ctypes.byref(g),
                                    ctypes.byref(l)))
    return dict(persistent=p.value == 1, plannable=p.value >= 1, grid=g.value, lds_bytes=l.value)


def bench_bsolve(cheb: Handle, reps: int = 10, fp32: bool = False):
    """(seconds per whole solve, algorithmic bytes per solve)."""
    s, b = _d(), _d()
    _check(_lib().nmg_bench_bsolve(_i64(cheb), ctypes.c_int(reps), ctypes.c_int(int(fp32)), ctypes.byref(s),
                                   ctypes.byref(b)))
    return s.value, b.value


def gemm_tn(A, B):
    """A^T B on the device by the tall-skinny kernels (A: n x m, B: n x k)."""
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    n, m = A.shape
    k = B.shape[1]
    C = np.empty((m, k), order="F")
    _check(_lib().nmg_gemm_tn(_i64(n), ctypes.c_int(m), ctypes.c_int(k), A.ctypes.data_as(_f64p),
                              B.ctypes.data_as(_f64p), C.ctypes.data_as(_f64p)))
    return C


def dense_product(op: int, kernel: int, A, B, C=None, alpha=1.0, beta=0.0):
    """op 0: alpha A B + beta C (A n x m, B m x K); op 1: A^T B (A n x m,
    B n x K).  kernel 1: FP64 MFMA kernels; 0: rocBLAS / FMA split reduction."""
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    n, m = A.shape
    K = B.shape[1]
    if op == 0:
        out = np.asfortranarray(np.zeros((n, K)) if C is None else C, dtype=np.float64).copy(order="F")
    else:
        out = np.zeros((m, K), order="F")
    _check(_lib().nmg_dense_product(ctypes.c_int(op), ctypes.c_int(kernel), _i64(n), ctypes.c_int(m),
                                    ctypes.c_int(K), _d(alpha), _p(A, _f64p), _p(B, _f64p), _d(beta),
                                    _p(out, _f64p)))
    return out


def bench_dense(op: int, kernel: int, n: int, m: int, K: int, reps: int = 10):
    """(seconds per product, algorithmic bytes, flops)."""
    sec, b, f = _d(), _d(), _d()
    _check(_lib().nmg_bench_dense(ctypes.c_int(op), ctypes.c_int(kernel), _i64(n), ctypes.c_int(m), ctypes.c_int(K),
                                  ctypes.c_int(reps), ctypes.byref(sec), ctypes.byref(b), ctypes.byref(f)))
    return sec.value, b.value, f.value


def set_dense_kernels(h: Handle, mode: int = 2):
    """2: MFMA except narrow A^T B (FMA split reduction); 1: MFMA throughout; 0: rocBLAS + FMA."""
    _check(_lib().nmg_set_dense_kernels(_i64(h), ctypes.c_int(int(mode))))


def bench_copy(n: int, reps: int = 10) -> float:
    """Mean seconds of a plain copy kernel of n doubles (16 n bytes)."""
    sec = _d()
    _check(_lib().nmg_bench_copy(_i64(n), ctypes.c_int(reps), ctypes.byref(sec)))
    return sec.value


def bench_gemm_tn(n: int, m: int, kc: int, reps: int = 10):
    """Mean seconds and algorithmic bytes of the re-orthogonalisation
    product C = A^T B (A n x m, B n x kc) on device arrays."""
    s, b = _d(), _d()
    _check(_lib().nmg_bench_gemm_tn(_i64(n), ctypes.c_int(m), ctypes.c_int(kc), ctypes.c_int(reps),
                                    ctypes.byref(s), ctypes.byref(b)))
    return s.value, b.value


def bench_spmv(mat: Handle, reps: int = 20, fp32: bool = False):
    s, b = _d(), _d()
    _check(_lib().nmg_bench_spmv(_i64(mat), ctypes.c_int(reps), ctypes.c_int(int(fp32)), ctypes.byref(s),
                                 ctypes.byref(b)))
    return s.value, b.value


def bench_block_chebstep(cheb: Handle, K: int, reps: int = 20, fp32: int = 0):
    """fp32: 0 FP64, 1 FP32 values, 2 FP32 values and vectors."""
    s, b = _d(), _d()
    _check(_lib().nmg_bench_block_chebstep(_i64(cheb), ctypes.c_int(K), ctypes.c_int(reps), ctypes.c_int(int(fp32)),
                                           ctypes.byref(s), ctypes.byref(b)))
    return s.value, b.value


def set_kernel_map(mat: Handle, group: int = 0, K: int = 0, rows_per_wave: int = 0):
    _check(_lib().nmg_set_kernel_map(_i64(mat), ctypes.c_int(group), ctypes.c_int(K), ctypes.c_int(rows_per_wave)))


def set_xcd_order(mat: Handle, on: bool = True):
    _check(_lib().nmg_set_xcd_order(_i64(mat), ctypes.c_int(int(on))))


def set_stream_loads(mat: Handle, on: bool = True):
    _check(_lib().nmg_set_stream_loads(_i64(mat), ctypes.c_int(int(on))))


def set_value_order(mat: Handle, tiled: bool = True):
    _check(_lib().nmg_set_value_order(_i64(mat), ctypes.c_int(int(tiled))))


def set_tiling(mat: Handle, K: int, max_nodes: int, rows_per_wave: int = 1, threads: int = 1024):
    _check(_lib().nmg_set_tiling(_i64(mat), ctypes.c_int(K), ctypes.c_int(max_nodes), ctypes.c_int(rows_per_wave),
                                 ctypes.c_int(threads)))


def set_tiled_wide(mat: Handle, K: int, max_nodes: int, lanes: int):
    """K-column scalar-mass products on LDS tiles of at most max_nodes nodes,
    one entry per lane (lanes per row 4, 8 or 16); max_nodes <= 0 drops it."""
    _check(_lib().nmg_set_tiled_wide(_i64(mat), ctypes.c_int(K), ctypes.c_int(max_nodes), ctypes.c_int(lanes)))


def set_wide(mat: Handle, K: int, lanes: int):
    """K-column scalar-mass products with one entry per lane (lanes per node
    row 4, 8, 16 or 32; 0 drops the layout)."""
    _check(_lib().nmg_set_wide(_i64(mat), ctypes.c_int(K), ctypes.c_int(lanes)))


def set_row_pipeline(mat: Handle, lanes: int, rows: int):
    """Single-column products of a 3x3-block matrix through pipelined lane
    groups: `lanes` lanes (16, 32, 64) walk `rows` consecutive node rows (2,
    4, 8), loading row j+1's entries before row j's gathers; rows=0 drops it."""
    _check(_lib().nmg_set_row_pipeline(_i64(mat), ctypes.c_int(lanes), ctypes.c_int(rows)))


def set_row_groups(mat: Handle, per_wave: int):
    """Single-column products through row groups (per_wave groups per
    wavefront: 1, 2, 4 or 8; 0 drops them)."""
    _check(_lib().nmg_set_row_groups(_i64(mat), ctypes.c_int(per_wave)))


def csr_autotune(mat: Handle, K: int, reps: int = 10, fp32: bool = False):
    s = _d()
    kind, rows, cap, thr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(_lib().nmg_csr_autotune(_i64(mat), ctypes.c_int(K), ctypes.c_int(reps), ctypes.c_int(int(fp32)),
                                   ctypes.byref(s), ctypes.byref(kind), ctypes.byref(rows), ctypes.byref(cap),
                                   ctypes.byref(thr)))
    return dict(seconds=s.value, kind=["lanes_per_row", "rows_per_wave", "tiled", "row_groups", "wide", "tiled_wide",
                                      "pipelined_rows"][kind.value],
                rows=rows.value, cap=cap.value, threads=thr.value)


def bench_chebstep(cheb: Handle, reps: int = 20, fp32: bool = False):
    s, b = _d(), _d()
    _check(_lib().nmg_bench_chebstep(_i64(cheb), ctypes.c_int(reps), ctypes.c_int(int(fp32)),
                                     ctypes.byref(s), ctypes.byref(b)))
    return s.value, b.value
