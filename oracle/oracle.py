"""ORACLE — test infrastructure only.

Python front of liboracle.so (oracle/nm_oracle.c), the CPU restatement of the
filtered-Lanczos path that the reference drives through pEVSL.  Only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import this
module, and only as the checker or as the timed CPU baseline.

Parity pins (tests/test_oracle_*.py):
  spmv, diagscale      brute force (scipy.sparse) and the closed form diag(B~) = 1
  chebiter             closed form b - B x = T_{k}((th-B)/de) b / T_{k}(th/de),
                       evaluated by dense eigendecomposition
  tridiag eig / sygv   scipy.linalg.eigh_tridiagonal / scipy.linalg.eigh
  findpol              p(gam) = 1, endpoint values vs threshold, minimal degree,
                       balance p(a) = p(b), numpy.polynomial.chebyshev values
  chebav               p applied to the eigenvalues of a dense pencil
  lanbounds            enclosure of the dense spectrum
  cheblannr            dense scipy.linalg.eigh(A, B) on the same pencil
                       (eigenvalues to 1e-12 relative, same count)
  kpm_count            quadrature of the damped density vs the closed-form
                       integral; exact moments of a known spectrum
  FluidPencil          the dense Schur complement D (Ad - E Ap^-1 ET) D of an
                       assembled fluid-solid body (operator) and p applied to
                       its dense eigenvalues (filter)
  shift_invert_eigs    closed-form eigenvalues of the 1D P1 finite-element
                       pencil (stiffness, consistent mass) at the low end of a
                       wide spectrum, where a dense QR-based solver is limited
                       to absolute errors of eps * ||A||
  FindPol's exact degree rule, ChebLanNr's test cadence and LanBounds'
  stopping rule are restated from the published EVSL/pEVSL design and are
  "parity unpinned" against pEVSL itself (not vendored, cannot be run).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SO = HERE / "liboracle.so"

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


class _CSR(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int64), ("n_cols", ctypes.c_int64), ("rowptr", _i64p),
                ("col", _i32p), ("val", _f64p)]


class _Prob(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("A", _CSR), ("d", _f64p), ("B", _CSR),
                ("blmin", ctypes.c_double), ("blmax", ctypes.c_double), ("bdeg", ctypes.c_int),
                ("amode", ctypes.c_int)]


class _Pol(ctypes.Structure):
    _fields_ = [("deg", ctypes.c_int), ("type", ctypes.c_int), ("cc", ctypes.c_double),
                ("dd", ctypes.c_double), ("gam", ctypes.c_double), ("bar", ctypes.c_double),
                ("aa", ctypes.c_double), ("bb", ctypes.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not SO.exists():
            build()
        L = ctypes.CDLL(str(SO))
        L.or_wtime.restype = ctypes.c_double
        L.or_diagscale.restype = ctypes.c_int64
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


class CSRView:
    """Keeps numpy arrays alive for a ctypes CSR struct."""

    def __init__(self, rowptr, col, val, n_cols=None):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        n = self.rowptr.size - 1
        self.n_rows = n
        self.n_cols = n if n_cols is None else n_cols
        self.s = _CSR(n, self.n_cols, _p(self.rowptr, _i64p),
                      _p(self.col, _i32p), _p(self.val, _f64p))

    @classmethod
    def of(cls, M):
        """From an object with rowptr/col/val (package CSR) or a scipy csr_matrix."""
        if isinstance(M, CSRView):
            return M
        if hasattr(M, "indptr"):
            return cls(M.indptr, M.indices, M.data, M.shape[1])
        return cls(M.rowptr, M.col, M.val, M.n_cols)


@dataclass
class Pol:
    deg: int
    type: int
    cc: float
    dd: float
    gam: float
    bar: float
    aa: float
    bb: float
    mu: np.ndarray


class Pencil:
    """The scaled pencil (A~ = D A D, B~ = D B D) with its Chebyshev B-solve.

    Mirrors what setupmatvec prepares (src/mod_matvec.f90:24-250): D from
    B's diagonal, B~ scaled, chebiter bounds from Lanczos on B~.
    """

    def __init__(self, A, B=None, *, bdeg=45, bounds_B=None, scale=True, seed=7):
        L = lib()
        self.A = CSRView.of(A)
        n = self.A.rowptr.size - 1
        self.n = n
        self.d = None
        self.Bt = None
        if B is not None:
            Bv = CSRView.of(B)
            val = Bv.val.copy()
            d = np.empty(n)
            if scale:
                rc = L.or_diagscale(ctypes.c_int64(n), _p(Bv.rowptr, _i64p), _p(Bv.col, _i32p),
                                    _p(val, _f64p), ctypes.c_double(1.0), _p(d, _f64p))
                if rc != 0:
                    raise ValueError(f"B diagonal not positive at row {-rc - 1}")
                self.d = d
            self.Bt = CSRView(Bv.rowptr, Bv.col, val, n)
        self.bdeg = bdeg
        self.blmin, self.blmax = (0.0, 0.0)
        self._mk()
        if B is not None:
            if bounds_B is None:
                bounds_B = self.lanbounds(geneig=False, mlan=1000, maxit=2000, tol=1e-12, seed=seed,
                                          on_B=True)[:2]
            self.blmin, self.blmax = bounds_B
            self._mk()

    def _mk(self):
        empty = _CSR(0, 0, None, None, None)
        self.s = _Prob(self.n, self.A.s, _p(self.d, _f64p), self.Bt.s if self.Bt else empty,
                       self.blmin, self.blmax, self.bdeg, 0)

    @property
    def geneig(self):
        return self.Bt is not None

    # --- operators -------------------------------------------------------
    def apply_a(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        w = np.empty(self.n)
        lib().or_apply_a(ctypes.byref(self.s), _p(x, _f64p), _p(y, _f64p), _p(w, _f64p))
        return y

    def apply_b(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        lib().or_apply_b(ctypes.byref(self.s), _p(x, _f64p), _p(y, _f64p))
        return y

    def bsolve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty(self.n)
        w = np.empty(3 * self.n)
        lib().or_bsolve(ctypes.byref(self.s), _p(b, _f64p), _p(x, _f64p), _p(w, _f64p))
        return x

    def chebav(self, pol: Pol, z):
        z = np.ascontiguousarray(z, dtype=np.float64)
        y = np.empty(self.n)
        w = np.empty(8 * self.n)
        lib().or_chebav(ctypes.byref(self.s), ctypes.c_int(int(self.geneig)), ctypes.c_int(pol.deg),
                        _p(pol.mu, _f64p), ctypes.c_double(pol.cc), ctypes.c_double(pol.dd),
                        _p(z, _f64p), _p(y, _f64p), _p(w, _f64p))
        return y

    # --- solvers -----------------------------------------------------------
    def lanbounds(self, *, geneig=None, mlan=3000, maxit=5000, tol=1e-5, seed=7, on_B=False):
        """Returns (lmin, lmax, nsteps) of B~^-1 A~ (or of B~ with on_B)."""
        if geneig is None:
            geneig = self.geneig
        s = _Prob.from_buffer_copy(self.s)
        if on_B:
            s.amode = 1
            s.B = self.Bt.s
            geneig = False
        lmin = ctypes.c_double()
        lmax = ctypes.c_double()
        ns = ctypes.c_int()
        lib().or_lanbounds(ctypes.byref(s), ctypes.c_int(int(geneig)), ctypes.c_int(mlan),
                           ctypes.c_int(maxit), ctypes.c_double(tol), ctypes.c_uint64(seed),
                           ctypes.byref(lmin), ctypes.byref(lmax), ctypes.byref(ns))
        return lmin.value, lmax.value, ns.value

    def cheblannr(self, xintv, pol: Pol, *, maxit=9624, tol=1e-5, rtol=1e-9, ntest=30, cycle=20,
                  seed=7, max_nev=None):
        xintv = np.ascontiguousarray(xintv, dtype=np.float64)
        max_nev = max_nev or min(self.n, 4000)
        maxit = min(maxit, self.n)
        lam = np.empty(max_nev)
        X = np.empty((max_nev, self.n))
        nev = ctypes.c_int()
        ns = ctypes.c_int()
        nc = ctypes.c_double()
        ps = _Pol(pol.deg, pol.type, pol.cc, pol.dd, pol.gam, pol.bar, pol.aa, pol.bb)
        lib().or_cheblannr(ctypes.byref(self.s), ctypes.c_int(int(self.geneig)), _p(xintv, _f64p),
                           ctypes.c_int(maxit), ctypes.c_double(tol), ctypes.c_double(rtol),
                           ctypes.c_int(ntest), ctypes.c_int(cycle), ctypes.c_int(pol.deg),
                           _p(pol.mu, _f64p), ctypes.byref(ps), ctypes.c_uint64(seed),
                           ctypes.c_int(max_nev), ctypes.byref(nev), _p(lam, _f64p), _p(X, _f64p),
                           ctypes.byref(ns), ctypes.byref(nc))
        k = nev.value
        return lam[:k].copy(), X[:k].copy(), ns.value, int(nc.value)

    def residuals(self, lam, X):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        X = np.ascontiguousarray(X, dtype=np.float64)
        r1 = np.empty(lam.size)
        r2 = np.empty(lam.size)
        lib().or_residuals(ctypes.byref(self.s), ctypes.c_int(lam.size), _p(lam, _f64p),
                           _p(X, _f64p), _p(r1, _f64p), _p(r2, _f64p))
        return r1, r2


class FluidPencil:
    """The fluid-solid pencil as setupmatvec prepares it for sparsefsAV
    (src/mod_matvec.f90:98-222, 498-520): D from diag(B), B~ = D B D with its
    Chebyshev solve, Dp from diag(-Ap), Ap~ = Dp (-Ap) Dp with its Chebyshev
    solve, and the operator
        A~ x = D [Ad (D x) + E (Dp C_p (Dp ET (D x)))],  C_p ~ Ap~^-1,
    i.e. D (Ad - E Ap^-1 ET) D.  Built from the C primitives (CSR product and
    chebiter); the filter recurrence is or_chebav's.  Test infrastructure and
    the bench's CPU baseline only."""

    def __init__(self, Q, *, bdeg, bounds_B, apdeg, bounds_Ap):
        self.n = Q.n
        self.d, bval = diagscale(Q.B, 1.0)
        self.Bt = CSRView(Q.B.rowptr, Q.B.col, bval)
        self.dp, apval = diagscale(Q.Ap, -1.0)
        self.Apt = CSRView(Q.Ap.rowptr, Q.Ap.col, apval)
        self.Ad, self.E, self.ET = (CSRView.of(M) for M in (Q.Ad, Q.E, Q.ET))
        self.bdeg, (self.blmin, self.blmax) = bdeg, bounds_B
        self.apdeg, (self.aplmin, self.aplmax) = apdeg, bounds_Ap

    def apply_a(self, x):
        v0 = self.d * x
        w0 = spmv(self.Ad, v0)
        x1 = self.dp * spmv(self.ET, v0)
        y1 = self.dp * chebiter(self.Apt, self.aplmin, self.aplmax, self.apdeg, x1)
        return self.d * (w0 + spmv(self.E, y1))

    def apply_b(self, x):
        return spmv(self.Bt, x)

    def bsolve(self, b):
        return chebiter(self.Bt, self.blmin, self.blmax, self.bdeg, b)

    def chebav(self, pol: Pol, z):
        """y = sum_k mu_k v_k, v_{k+1} = t (A~ B~^-1 v_k - cc v_k) - v_{k-1}."""
        vk = np.array(z, dtype=np.float64)
        vkm = np.zeros_like(vk)
        y = pol.mu[0] * vk
        for k in range(1, pol.deg + 1):
            t = (1.0 if k == 1 else 2.0) / pol.dd
            v = t * (self.apply_a(self.bsolve(vk)) - pol.cc * vk) - (0.0 if k == 1 else vkm)
            y += pol.mu[k] * v
            vkm, vk = vk, v
        return y


def findpol(xintv, thr_int=0.8, thr_ext=0.7, *, min_deg=2, max_deg=10000, damping=2) -> Pol:
    xintv = np.ascontiguousarray(xintv, dtype=np.float64)
    mu = np.zeros(max_deg + 1)
    ps = _Pol()
    rc = lib().or_findpol(_p(xintv, _f64p), ctypes.c_double(thr_int), ctypes.c_double(thr_ext),
                          ctypes.c_int(min_deg), ctypes.c_int(max_deg), ctypes.c_int(damping),
                          _p(mu, _f64p), ctypes.byref(ps))
    if rc != 0:
        raise ValueError(f"findpol failed ({rc})")
    return Pol(ps.deg, ps.type, ps.cc, ps.dd, ps.gam, ps.bar, ps.aa, ps.bb, mu[:ps.deg + 1].copy())


def chebiter(B, lmin, lmax, deg, b):
    Bv = CSRView.of(B)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.empty(b.size)
    w = np.empty(3 * b.size)
    lib().or_chebiter(ctypes.byref(Bv.s), ctypes.c_double(lmin), ctypes.c_double(lmax), ctypes.c_int(deg),
                      _p(b, _f64p), _p(x, _f64p), _p(w, _f64p))
    return x


def spmv(M, x):
    Mv = CSRView.of(M)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(Mv.rowptr.size - 1)
    lib().or_spmv(ctypes.byref(Mv.s), _p(x, _f64p), _p(y, _f64p))
    return y


def diagscale(M, sign=1.0):
    Mv = CSRView.of(M)
    n = Mv.rowptr.size - 1
    val = Mv.val.copy()
    d = np.empty(n)
    rc = lib().or_diagscale(ctypes.c_int64(n), _p(Mv.rowptr, _i64p), _p(Mv.col, _i32p), _p(val, _f64p),
                            ctypes.c_double(sign), _p(d, _f64p))
    if rc != 0:
        raise ValueError(f"diagonal not positive at row {-rc - 1}")
    return d, val


def tridiag_eig(d, e):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    m = d.size
    w = np.empty(m)
    Z = np.empty((m, m))
    rc = lib().or_tridiag_eig(ctypes.c_int(m), _p(d, _f64p), _p(e, _f64p), _p(w, _f64p), _p(Z, _f64p))
    if rc != 0:
        raise RuntimeError("tridiagonal QR did not converge")
    return w, Z.T.copy()  # columns = eigenvectors


def tridiag_eigvals(d, e):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    w = np.empty(d.size)
    lib().or_tridiag_eigvals(ctypes.c_int(d.size), _p(d, _f64p), _p(e, _f64p), _p(w, _f64p))
    return w


def sygv(GA, GB):
    GA = np.ascontiguousarray(GA, dtype=np.float64)
    GB = np.ascontiguousarray(GB, dtype=np.float64)
    n = GA.shape[0]
    mu = np.empty(n)
    W = np.empty((n, n))
    rc = lib().or_sygv(ctypes.c_int(n), _p(GA, _f64p), _p(GB, _f64p), _p(mu, _f64p), _p(W, _f64p))
    if rc != 0:
        raise ValueError("G_B not positive definite")
    return mu, W.T.copy()


def start_vector(n, seed=7):
    v = np.empty(n)
    lib().or_start_vector(ctypes.c_int64(n), ctypes.c_uint64(seed), _p(v, _f64p))
    return v


def set_threads(t):
    lib().or_set_threads(ctypes.c_int(int(t)))


def num_threads():
    return int(lib().or_num_threads())


def wtime():
    return lib().or_wtime()


def set_verbose(every: int):
    """Progress lines on stderr every ``every`` Lanczos steps of the C
    restatement (0: silent), for long CPU solves."""
    lib().or_set_verbose(ctypes.c_int(int(every)))


def solve_interval(A, B, xintv_ab, *, bdeg=45, thr_int=0.8, thr_ext=0.7, lb=(3000, 5000, 1e-5),
                   seed=7, tol=1e-5, rtol=1e-9, maxit=9624, return_info=False):
    """Full restated path of pnm_apply_pevsl (src/mod_pevsl.f90:54-162) on CPU.

    ``xintv_ab`` = (a, b) in lambda units; a < 1e-10 means "from LMIN"
    (src/mod_pevsl.f90:98-100).
    """
    P = Pencil(A, B, bdeg=bdeg, seed=seed)
    lmin, lmax, nlb = P.lanbounds(mlan=lb[0], maxit=lb[1], tol=lb[2], seed=seed)
    a, b = xintv_ab
    if a < 1e-10:
        a = lmin
    xintv = np.array([a, b, lmin, lmax])
    pol = findpol(xintv, thr_int, thr_ext)
    lam, X, nst, ncand = P.cheblannr(xintv, pol, maxit=maxit, tol=tol, rtol=rtol, seed=seed)
    r1, r2 = P.residuals(lam, X)
    info = dict(lmin=lmin, lmax=lmax, blmin=P.blmin, blmax=P.blmax, lanbounds_steps=nlb,
                deg=pol.deg, bar=pol.bar, nsteps=nst, ncand=ncand, res_ref=r1, res2=r2, pencil=P,
                xintv=xintv)
    if return_info:
        return lam, X, info
    return lam, X


def jackson(mdeg):
    """Jackson damping factors g_k, k = 0..mdeg (Weisse et al., Rev. Mod.
    Phys. 78, 275 (2006), eq. 71, with N = mdeg + 1 moments)."""
    N = mdeg + 1
    k = np.arange(N)
    a = np.pi / (N + 1)
    return ((N - k + 1) * np.cos(k * a) + np.sin(k * a) / np.tan(a)) / (N + 1)


def kpm_count(tau, lmin, lmax, a, b, npts=200001):
    """Jackson-damped KPM eigenvalue count in [a, b] by trapezoidal quadrature
    of the density in theta (t = cos theta), independent of a closed-form
    integral:  count = 1/pi int_{theta_b}^{theta_a} sum_k (2-[k=0]) g_k tau_k cos(k theta) dtheta."""
    tau = np.asarray(tau, dtype=float)
    cc, dd = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    th = lambda x: np.arccos(np.clip((x - cc) / dd, -1.0, 1.0))
    ta, tb = th(a), th(b)
    t = np.linspace(tb, ta, npts)
    g = jackson(tau.size - 1)
    w = g * tau
    w[1:] *= 2.0
    dens = np.polynomial.chebyshev.chebval(np.cos(t), w)  # sum_k w_k cos(k t) = sum_k w_k T_k(cos t)
    return float(np.trapezoid(dens, t) / np.pi)


def rademacher_block(n, K, seed, batch):
    """The K Rademacher vectors of batch `batch` used for the KPM trace
    estimate (splitmix64 stream of the Lanczos start vector, signs), as an
    (n, K) array."""
    s = (seed ^ ((0xD1B54A32D192ED03 * (batch + 1)) & 0xFFFFFFFFFFFFFFFF)) & 0xFFFFFFFFFFFFFFFF
    r = start_vector(n * K, s)
    return np.where(r < 0.0, -1.0, 1.0).reshape(n, K)


def kpm_moments_dense(At, Bt, lmin, lmax, mdeg, V):
    """tau_k = mean_c v_c^T T_k((At Bt^-1 - cc)/dd) v_c with a dense solve."""
    import scipy.linalg as sl
    cc, dd = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    C = sl.cho_factor(Bt)
    op = lambda X: (At @ sl.cho_solve(C, X) - cc * X) / dd
    tau = np.empty(mdeg + 1)
    Tm, T = V, op(V)
    tau[0] = np.sum(V * V) / V.shape[1]
    tau[1] = np.sum(V * T) / V.shape[1]
    for k in range(2, mdeg + 1):
        Tm, T = T, 2.0 * op(T) - Tm
        tau[k] = np.sum(V * T) / V.shape[1]
    return tau


os.environ.setdefault("OMP_PROC_BIND", "false")


def shift_invert_eigs(A, B, a, b, *, extra=8, sigma=None):
    """Eigenvalues of the symmetric pencil (A, B) in [a, b] by shift-invert
    Lanczos (ARPACK) at sigma (default: the interval centre), for dense or
    sparse A, B.  The eigenvalues nu = 1/(lam - sigma) are computed to
    machine precision relative to the largest |nu|, so lam near the interval
    carries errors ~eps |lam - sigma|, independent of ||A||.  A dense
    QR-based solver (LAPACK sygvd) instead leaves absolute errors ~eps ||A||,
    which at the bottom of a wide spectrum exceed 1e-12 relative.
    The count inside [a, b] is found by enlarging k until an eigenvalue
    outside the interval is reached on both sides."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    sigma = 0.5 * (a + b) if sigma is None else sigma
    n = A.shape[0]
    dense = not sp.issparse(A)
    k = extra
    while True:
        k = min(k, n - 1)
        if dense:
            w = spl.eigsh(np.asarray(A), k=k, M=np.asarray(B), sigma=sigma, which="LM", tol=0.0,
                          return_eigenvectors=False)
        else:
            w = spl.eigsh(A.tocsc(), k=k, M=B.tocsc(), sigma=sigma, which="LM", tol=0.0,
                          return_eigenvectors=False)
        w = np.sort(w)
        if (w[0] < a and w[-1] > b) or k == n - 1:
            return w[(w >= a) & (w <= b)]
        k = 2 * k
