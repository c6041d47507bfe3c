"""The reference's operator set-up and eigensolver driver on the nmgpu C-ABI.

``setupmatvec`` mirrors src/mod_matvec.f90:24-250 (solid case) and
``apply_pevsl`` mirrors pnm_apply_pevsl, src/mod_pevsl.f90:16-222: same call
order, same parameters (MLAN 3000, LANSTEP 5000, TOL 1e-5, thresholds
0.8/0.7, MAXIT = 3*max(4*(EVINT+2), 2000), B-bounds (1000, 2000, 1e-12),
chebiter degree 45 for P2 and 25 for P1).  The interval may be given in mHz
(as the reference's lowfreq/upfreq) or directly in lambda = omega^2.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import pevsl as pv
from .assemble import FluidProblem

__all__ = ["MatvecSetup", "FluidSetup", "FluidProblem", "setupmatvec_fs", "ap_chebiter_degree", "upload", "setup_bsolve", "setupmatvec", "apply_pevsl", "apply_slices", "plan_slice_waves", "refine_bsolves", "refine_inner_degree", "kpm_degree", "solve_input", "EigResult", "chebiter_degree",
           "freq_to_lambda", "lambda_to_freq"]

EVINT = 800


def freq_to_lambda(f_mhz: float) -> float:
    """lambda = (2 pi f)^2 * 1e-6 for f in mHz (src/mod_pevsl.f90:93-94), with
    the reference's roundings: f is a default REAL (src/mod_para.f90:51-52)
    and PI is assigned from a default-real literal (src/mod_pevsl.f90:43),
    so both enter the double-precision product as float32 values."""
    pi32 = float(np.float32(3.14159265359))
    return (2.0 * pi32 * float(np.float32(f_mhz))) ** 2 * 1.0e-6


def lambda_to_freq(lam):
    """f [mHz] = sqrt(lambda)/(2 pi) * 1e3 (src/mod_pevsl.f90:177-181)."""
    return np.sqrt(np.maximum(lam, 0.0)) / (2.0 * math.pi) * 1.0e3


def chebiter_degree(porder: int, blmin: float, blmax: float, policy="reference", eps=1e-13) -> int:
    """Degree of the Chebyshev B-solve.

    ``reference``: 45 for P2, 25 for P1 (src/mod_matvec.f90:88-92).
    ``auto``: the smallest c with 2 rho^c <= eps, rho = (sqrt k - 1)/(sqrt k + 1),
    k = blmax/blmin -- enough for the B-solve error to stay below eps.
    An integer selects that degree.
    """
    if isinstance(policy, int):
        return policy
    ref = 45 if porder == 2 else 25
    if policy == "reference":
        return ref
    if policy == "auto":
        k = blmax / blmin
        rho = (math.sqrt(k) - 1.0) / (math.sqrt(k) + 1.0)
        return max(1, int(math.ceil(math.log(eps / 2.0) / math.log(rho))))
    raise ValueError(policy)


def refine_inner_degree(lmin: float, lmax: float, per_pass: float = 1e-6) -> int:
    """Inner degree of a refinement pass: the smallest c with 2 rho^c <= per_pass
    (below that the FP32 iterates' own floor, ~1e-6, dominates)."""
    k = lmax / lmin
    rho = (math.sqrt(k) - 1.0) / (math.sqrt(k) + 1.0)
    return max(1, int(math.ceil(math.log(per_pass / 2.0) / math.log(rho))))


def refine_bsolves(mv: "MatvecSetup", passes: int = 2, per_pass: float = 1e-6, ap: bool = True) -> dict:
    """Switch the B~ solve (and the Ap~ solve of a fluid-solid set-up) to
    mixed-precision refinement: ``passes`` FP32 inner solves on FP64
    residuals (``pevsl.chebiter_set_refine``); passes = 0 restores the plain
    FP64 solves.  Returns the inner degrees."""
    out = {}
    deg = refine_inner_degree(mv.blmin, mv.blmax, per_pass) if passes else 0
    pv.chebiter_set_refine(mv.chebB, passes, deg)
    out["B"] = deg
    if ap and mv.fs is not None:
        dp = refine_inner_degree(mv.fs.aplmin, mv.fs.aplmax, per_pass) if passes else 0
        pv.chebiter_set_refine(mv.fs.chebAp, passes, dp)
        out["Ap"] = dp
    return out


@dataclass
class MatvecSetup:
    """Device-side counterpart of the reference's ``mymatvec``
    (src/mod_cg_datatype.f90:88-115)."""
    n: int
    matA: int
    matB: int
    d: np.ndarray
    blmin: float
    blmax: float
    bdeg: int
    chebB: int
    porder: int
    timings: dict = field(default_factory=dict)
    tuned: dict = field(default_factory=dict)
    fs: "FluidSetup | None" = None   # fluid-solid operator (A unused when set)

    def free(self):
        if self.chebB:
            pv.pevsl_chebiter_free(self.chebB)
        if self.matA:
            pv.csr_free(self.matA)
        pv.csr_free(self.matB)
        if self.fs is not None:
            self.fs.free()


@dataclass
class FluidSetup:
    """The fluid-solid handles of ``mymatvec`` (sAdV, sEV, sETV, sApV, chebAp;
    src/mod_cg_datatype.f90:92-96) and the pressure scaling Dp."""
    matAd: int
    matE: int
    matET: int
    matAp: int
    dp: np.ndarray
    aplmin: float = 0.0
    aplmax: float = 0.0
    apdeg: int = 0
    chebAp: int = 0

    def free(self):
        if self.chebAp:
            pv.pevsl_chebiter_free(self.chebAp)
        for m in (self.matAd, self.matE, self.matET, self.matAp):
            pv.csr_free(m)


def ap_chebiter_degree(porder: int, lmin: float, lmax: float, policy="reference", eps=1e-13) -> int:
    """Degree of the Chebyshev solve with Ap~.

    ``reference`` (src/mod_matvec.f90:166-171): P2 max(min(int(lmax/lmin),
    15)*3, 100), P1 25.  ``auto``: the smallest degree whose error bound
    2 rho^c stays below eps (as for the B-solve).  An integer selects it."""
    if isinstance(policy, int):
        return policy
    if policy == "auto":
        return chebiter_degree(porder, lmin, lmax, "auto", eps)
    if policy != "reference":
        raise ValueError(policy)
    if porder == 2:
        return max(min(int(lmax / lmin), 15) * 3, 100)
    return 25


def setupmatvec_fs(problem: FluidProblem, *, bdeg="reference", seed: int = 7, bounds=(1000, 2000, 1e-12),
                   ap_bounds=(2000, 3000, 1e-12), tune=(), apdeg="reference", verbose: bool = False) -> MatvecSetup:
    """setupmatvec for a body with fluid regions (src/mod_matvec.f90:24-222):
    B scaled and its Chebyshev solve as in the solid case; Ad, E, ET uploaded;
    Ap negated and scaled (Apdiagscaling), its bounds by LanBounds and its
    Chebyshev solve (degree by ``apdeg``, see ``ap_chebiter_degree``).
    ``tune``: block widths for which the layouts of B and Ad are autotuned on
    the device.  ``verbose``: a progress line on stderr after each phase
    (long set-ups of large bodies)."""
    import sys as _sys
    t0 = time.perf_counter()

    def say(what):
        if verbose:
            print(f"[setupmatvec_fs] {what} at {time.perf_counter() - t0:.1f}s", file=_sys.stderr, flush=True)
    n = problem.n
    matB = pv.csr_create(problem.B.rowptr, problem.B.col, problem.B.val)
    d = pv.csr_diagscale(matB, 1.0)                                   # :47-71
    fs = FluidSetup(matAd=pv.csr_create(problem.Ad.rowptr, problem.Ad.col, problem.Ad.val),        # :98-119
                    matE=pv.csr_create(problem.E.rowptr, problem.E.col, problem.E.val, n_cols=problem.n_p),
                    matET=pv.csr_create(problem.ET.rowptr, problem.ET.col, problem.ET.val, n_cols=n),
                    matAp=pv.csr_create(problem.Ap.rowptr, problem.Ap.col, problem.Ap.val), dp=None)
    fs.dp = pv.csr_diagscale(fs.matAp, -1.0)                          # :137-140, 345-441
    say("uploaded")
    tuned = {}
    for K in tune:
        for key, mat in (("B", matB), ("Ad", fs.matAd)):
            if pv.csr_info(mat)["layout"] != "csr":
                tuned[f"{key}{K}"] = pv.csr_autotune(mat, K, 10)
        say(f"autotuned {K} columns")
    hp = pv.pevsl_start(0)                                            # :152
    try:
        pv.pevsl_setprobsizes(hp, problem.n_p, problem.n_p, -1)
        pv.pevsl_set_seed(hp, seed)
        pv.pevsl_setamv(hp, fs.matAp)                                 # :157
        if verbose:
            pv.set_verbose(hp, 100)
        fs.aplmin, fs.aplmax = pv.pevsl_lanbounds(hp, *ap_bounds)     # :160-162
    finally:
        pv.pevsl_finish(hp)
    say("Ap bounds")
    fs.apdeg = ap_chebiter_degree(problem.porder, fs.aplmin, fs.aplmax, apdeg)
    fs.chebAp = pv.pevsl_setup_chebiter(fs.aplmin, fs.aplmax, fs.apdeg, fs.matAp)   # :174
    mv = MatvecSetup(n=n, matA=0, matB=matB, d=d, blmin=0.0, blmax=0.0, bdeg=0, chebB=0,
                     porder=problem.porder, timings=dict(upload=time.perf_counter() - t0), fs=fs)
    mv.tuned = tuned
    mv = setup_bsolve(mv, bdeg=bdeg, seed=seed, bounds=bounds, verbose=verbose)
    say("B bounds and Chebyshev solve")
    return mv


def upload(problem, *, porder: int | None = None, tune=(1,)) -> MatvecSetup:
    """Copy A and B to HBM once and scale B in place (src/mod_matvec.f90:47-71,
    252-342).  Returns a set-up without the B-solve (see ``setup_bsolve``).

    ``tune``: block widths for which the product layouts of A and B are
    autotuned on the device (empty to keep the defaults)."""
    porder = porder or problem.mesh.porder
    t0 = time.perf_counter()
    n = problem.n
    matB = pv.csr_create(problem.B.rowptr, problem.B.col, problem.B.val)
    matA = pv.csr_create(problem.A.rowptr, problem.A.col, problem.A.val)
    d = pv.csr_diagscale(matB, 1.0)                      # Bdiagscaling, :252-342
    if pv.csr_info(matA)["layout"] == "bsr3":
        pv.set_value_order(matA, True)                   # 64-block value tiles (measured ~1.5% faster)
    t = dict(upload=time.perf_counter() - t0)
    t0 = time.perf_counter()
    tuned = {}
    for K in tune:
        for key, mat in (("B", matB), ("A", matA)):
            if pv.csr_info(mat)["layout"] != "csr":
                tuned[f"{key}{K}"] = pv.csr_autotune(mat, K, 10)
    t["autotune"] = time.perf_counter() - t0
    mv = MatvecSetup(n=n, matA=matA, matB=matB, d=d, blmin=0.0, blmax=0.0, bdeg=0, chebB=0,
                     porder=porder, timings=t)
    mv.tuned = tuned
    return mv


def setup_bsolve(mv: MatvecSetup, *, bdeg="reference", seed: int = 7,
                 bounds=(1000, 2000, 1e-12), verbose: bool = False) -> MatvecSetup:
    """Lanczos bounds of B~ and its Chebyshev solve (src/mod_matvec.f90:75-95)."""
    t0 = time.perf_counter()
    if mv.chebB:
        pv.pevsl_chebiter_free(mv.chebB)
        mv.chebB = 0
    hB = pv.pevsl_start(0)                               # :75
    try:
        pv.pevsl_setprobsizes(hB, mv.n, mv.n, -1)
        pv.pevsl_set_seed(hB, seed)
        pv.pevsl_setamv(hB, mv.matB)                     # :80
        if verbose:
            pv.set_verbose(hB, 100)
        mv.blmin, mv.blmax = pv.pevsl_lanbounds(hB, *bounds)   # :83-85
    finally:
        pv.pevsl_finish(hB)                              # :95
    mv.bdeg = chebiter_degree(mv.porder, mv.blmin, mv.blmax, bdeg)
    mv.chebB = pv.pevsl_setup_chebiter(mv.blmin, mv.blmax, mv.bdeg, mv.matB)   # :93
    mv.timings["setup_B"] = time.perf_counter() - t0
    return mv


def setupmatvec(problem, *, porder: int | None = None, bdeg="reference", seed: int = 7,
                bounds=(1000, 2000, 1e-12), tune=(1,)) -> MatvecSetup:
    """Upload A and B, scale B, bound its spectrum and set up its Chebyshev
    solve (src/mod_matvec.f90:24-95, 225-247)."""
    mv = upload(problem, porder=porder, tune=tune)
    return setup_bsolve(mv, bdeg=bdeg, seed=seed, bounds=bounds)


@dataclass
class EigResult:
    lam: np.ndarray              # ascending
    X: np.ndarray | None         # (nev, n) scaled-coordinate eigenvectors x~
    res_ref: np.ndarray          # reference residual definition
    res2: np.ndarray             # ||A~x - lam B~x|| / |lam|, x B~-normalised
    lmin: float
    lmax: float
    xintv: np.ndarray
    deg: int
    bar: float
    lan_steps: int
    filter_applies: int
    ncand: int
    timings: dict
    lan_steps_bounds: int = 0
    probe: dict = field(default_factory=dict)
    handle: int = 0   # solver handle kept alive with keep=True (caller finishes it)
    slices: np.ndarray | None = None        # slice boundaries (apply_slices)
    slice_counts: np.ndarray | None = None  # pairs per slice
    dos_count: float | None = None          # KPM estimate of the count in the interval
    target_met: bool = True                 # every pair met the residual target at the stop
    waves: list | None = None               # slice ranges solved together (apply_slices with a budget)

    @property
    def freq_mhz(self):
        return lambda_to_freq(self.lam)

    def physical_vectors(self, d):
        """x = D x~ as saved by the reference (src/mod_pevsl.f90:199)."""
        return None if self.X is None else self.X * d[None, :]


def split_parts(block: int, split="auto") -> int:
    """Streams the k-column B-solve is split over: 'auto' picks 2 for blocks
    of 8 or more (two 4-column solves overlap their dispatch gaps), else 1."""
    if split == "auto":
        return 2 if block >= 8 and block % 2 == 0 else 0
    return int(split or 0)


def _open_solver(mv: MatvecSetup, *, precision, seed, rtol, probe_every=0, block=1, resid_target=None,
                 split="auto", graphs=False, verbose=0, split_filter=True, dense_kernels=2):
    """pEVSL_Start and the operator settings of src/mod_pevsl.f90:54-82."""
    h = pv.pevsl_start(0)                                     # :54
    try:
        pv.pevsl_setprobsizes(h, mv.n, mv.n, -1)              # :57
        pv.pevsl_set_seed(h, seed)
        pv.pevsl_setbmv(h, mv.matB)                           # :69
        pv.pevsl_setbsol_chebiter(h, 2, mv.chebB)             # :72-73
        if mv.fs is not None:                                 # :77-80
            pv.pevsl_setamv_fs(h, mv.fs.matAd, mv.fs.matE, mv.fs.matET, mv.fs.chebAp, mv.d, mv.fs.dp)
        else:
            pv.pevsl_setamv_dad(h, mv.matA, mv.d)
        pv.pevsl_set_geneig(h)                                # :82
        if precision != pv.FP64:
            pv.pevsl_set_precision(h, precision)
        pv.set_lanczos_controls(h, rtol=rtol)
        if probe_every > 0:
            pv.set_kernel_probe(h, probe_every)
        if block > 1:
            pv.set_block_size(h, block)
        parts = split_parts(block, split)
        if parts > 1:
            pv.set_split_streams(h, parts)
            if not split_filter:
                pv.set_split_filter(h, False)
        if graphs:
            pv.set_graphs(h, True)
        if dense_kernels != 2:
            pv.set_dense_kernels(h, dense_kernels)
        if verbose:
            pv.set_verbose(h, verbose)
        if resid_target is None and precision != pv.FP64:
            # the FP32 B-solve bounds ||A~x - lam B~x||/|lam| near 1e-6 (its
            # eigenvalues still agree with FP64 to ~1e-14): stop at that floor
            resid_target = 1e-5
        if resid_target is not None:
            pv.set_residual_target(h, resid_target)
    except BaseException:
        pv.pevsl_finish(h)
        raise
    return h


def apply_pevsl(mv: MatvecSetup, *, lowfreq: float | None = None, upfreq: float | None = None,
                interval=None, tol: float = 1e-5, thr_int: float = 0.8, thr_ext: float = 0.7,
                lanbounds=(3000, 5000, 1e-5), precision: int = pv.FP64, seed: int = 7,
                rtol: float = 1e-9, vectors: bool = True, residual_check: bool = True,
                probe_every: int = 0, block: int = 1, resid_target: float | None = None,
                keep: bool = False, split="auto", graphs: bool = False, verbose: int = 0,
                split_filter: bool = True, dense_kernels: int = 2) -> EigResult:
    """pnm_apply_pevsl (src/mod_pevsl.f90:16-222) on the device-resident operator.
    ``dense_kernels``: 2 the FP64 MFMA kernels for the block recurrence and the
    extraction's dense products except A^T B narrower than 16 columns (FMA
    split reduction), 1 MFMA throughout, 0 rocBLAS / the FMA split reduction."""
    if interval is None:
        a, b = freq_to_lambda(lowfreq), freq_to_lambda(upfreq)
    else:
        a, b = interval
    t = {}
    t0 = time.perf_counter()
    pol_live = False
    h = _open_solver(mv, precision=precision, seed=seed, rtol=rtol, probe_every=probe_every, block=block,
                     resid_target=resid_target, split=split, graphs=graphs, verbose=verbose,
                     split_filter=split_filter, dense_kernels=dense_kernels)
    try:
        t["setup"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        lmin, lmax = pv.pevsl_lanbounds(h, *lanbounds)        # :84
        t["lanbounds"] = time.perf_counter() - t0
        lb_steps = pv.solve_stats(h)["lan_steps"]
        if a < 1.0e-10:                                       # :98-100
            a = lmin
        xintv = np.array([a, b, lmin, lmax])
        t0 = time.perf_counter()
        pol = pv.pevsl_findpol(xintv, thr_int, thr_ext)       # :115
        pol_live = True
        pinfo = pv.pol_info(pol)
        t["findpol"] = time.perf_counter() - t0
        nev = EVINT + 2
        mlan = max(4 * nev, 2000)
        maxit = 3 * mlan                                      # :116-119
        t0 = time.perf_counter()
        pv.pevsl_cheblannr(h, xintv, maxit, tol, pol)         # :122
        t["cheblannr"] = time.perf_counter() - t0
        st = pv.solve_stats(h)
        target_met, _ = pv.solve_quality(h)
        probe = {}
        if probe_every > 0:
            for key, which in (("chebstep", 0), ("filterstep", 1)):
                sec, ns, by = pv.kernel_probe(h, which)
                probe[key] = dict(seconds=sec, samples=ns, bytes=by)
        t0 = time.perf_counter()
        if residual_check:
            r1, r2 = pv.residuals(h)                          # :137-162
        else:
            k = pv.pevsl_get_nev(h)
            r1 = r2 = np.full(k, np.nan)
        t["residual"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        lam, X = pv.pevsl_copy_result(h, mv.n, vectors)       # :124-130
        t["copy_result"] = time.perf_counter() - t0
        pol_live = False
        pv.pevsl_freepol(pol)                                 # :215
    except BaseException:
        if pol_live:
            pv.pevsl_freepol(pol)
        pv.pevsl_finish(h)
        raise
    if not keep:
        pv.pevsl_finish(h)                                    # :220
    order = np.argsort(lam, kind="stable")                    # ssort_real, :165
    return EigResult(lam=lam[order], X=None if X is None else X[order], res_ref=r1[order],
                     res2=r2[order], lmin=lmin, lmax=lmax, xintv=xintv, deg=pinfo.deg,
                     bar=pinfo.bar, lan_steps=st["lan_steps"], filter_applies=st["filter_applies"],
                     ncand=st["ncand"], timings=t, lan_steps_bounds=lb_steps, probe=probe,
                     handle=h if keep else 0, target_met=target_met)


def kpm_degree(xintv, n_slices: int, lo: int = 50, hi: int = 3000) -> int:
    """Degree of the KPM expansion that resolves n_slices slices of [a, b]:
    the Jackson kernel is ~pi/M wide in theta = arccos((x - cc)/dd), so M is
    chosen to give about two kernel widths per slice."""
    a, b, lmin, lmax = xintv
    cc, dd = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    th = lambda x: math.acos(min(1.0, max(-1.0, (x - cc) / dd)))
    dth = abs(th(a) - th(b))
    return int(min(hi, max(lo, math.ceil(2.0 * math.pi * n_slices / max(dth, 1e-12)))))


def plan_slice_waves(counts, n: int, budget_bytes: float, *, vectors_per_pair: float = 3.0, extra: int = 40,
                     max_width: int = 64):
    """Consecutive slices grouped into waves whose Lanczos bases fit a memory
    budget.  A slice holding c pairs is taken to build about
    vectors_per_pair * c + extra basis vectors, each kept twice (V and
    Z = B~V, FP64: 16 n bytes).  A wave holds at most ``max_width`` slices;
    a slice whose own basis exceeds the budget forms a wave by itself.
    Returns [(first, last + 1), ...] over the slice indices."""
    per = 16.0 * n
    need = [(vectors_per_pair * float(c) + extra) * per for c in counts]
    waves, i = [], 0
    while i < len(need):
        j, tot = i, 0.0
        while j < len(need) and j - i < max_width and (j == i or tot + need[j] <= budget_bytes):
            tot += need[j]
            j += 1
        waves.append((i, j))
        i = j
    return waves


def apply_slices(mv: MatvecSetup, *, interval, n_slices: int, slicer: str = "kpm", kpm_deg: int | None = None,
                 kpm_nvec: int = 16, tol: float = 1e-5, thr_int: float = 0.8, thr_ext: float = 0.7,
                 lanbounds=(3000, 5000, 1e-5), precision: int = pv.FP64, seed: int = 7, rtol: float = 1e-9,
                 vectors: bool = True, residual_check: bool = True, probe_every: int = 0,
                 resid_target: float | None = None, keep: bool = False, split="auto",
                 verbose: int = 0, basis_budget: float | None = None) -> EigResult:
    """The interval split into n_slices spectrum slices, solved in lockstep.

    Slice boundaries: ``kpm`` places them at equal counts of the kernel
    polynomial density of states (pEVSL kpmdos + spslicer); ``uniform``
    splits the interval evenly.  Each slice runs ChebLanNr with its own
    FindPol filter; every sparse product is shared by the slices of a wave.

    ``basis_budget`` (bytes): the slices run in waves of consecutive slices
    whose estimated Lanczos bases fit it (``plan_slice_waves``, counts from
    the KPM density of states); without it all slices form one wave.
    """
    a, b = interval
    t = {}
    t0 = time.perf_counter()
    width = 1 << max(0, (n_slices - 1).bit_length())   # the k-column width the slices run at
    h = _open_solver(mv, precision=precision, seed=seed, rtol=rtol, probe_every=probe_every,
                     resid_target=resid_target, split=split_parts(width, split), verbose=verbose)
    try:
        t["setup"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        lmin, lmax = pv.pevsl_lanbounds(h, *lanbounds)
        t["lanbounds"] = time.perf_counter() - t0
        lb_steps = pv.solve_stats(h)["lan_steps"]
        if a < 1.0e-10:
            a = lmin
        xintv = np.array([a, b, lmin, lmax])
        t0 = time.perf_counter()
        ecnt = None
        mu = None
        if slicer == "kpm" or basis_budget is not None:
            mdeg = kpm_deg or kpm_degree(xintv, n_slices)
            mu, ecnt = pv.kpmdos(h, mdeg, kpm_nvec, xintv)
        if slicer == "kpm":
            sli = pv.spslicer(mu, xintv, n_slices)
        elif slicer == "uniform":
            sli = np.linspace(a, b, n_slices + 1)
        else:
            raise ValueError(slicer)
        waves = [(0, n_slices)]
        if basis_budget is not None:
            est = [max(0.0, pv.dos_count(mu, [sli[i], sli[i + 1], lmin, lmax])) for i in range(n_slices)]
            waves = plan_slice_waves(est, mv.n, basis_budget)
        t["slicer"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        nev = EVINT + 2
        maxit = 3 * max(4 * nev, 2000)
        lams, Xs, r1s, r2s, counts = [], [], [], [], []
        steps = applies = ncand = 0
        target_met = True
        probe = {}
        t["residual"] = 0.0
        for (i0, i1) in waves:
            if len(waves) > 1:
                wv = 1 << max(0, (i1 - i0 - 1).bit_length())
                pv.set_split_streams(h, split_parts(wv, split))
            pv.solve_slices(h, sli[i0:i1 + 1], lmin, lmax, thr_int, thr_ext, maxit, tol)
            st = pv.solve_stats(h)
            steps += st["lan_steps"]
            applies += st["filter_applies"]
            ncand += st["ncand"]
            target_met = target_met and pv.solve_quality(h)[0]
            counts.append(pv.slice_counts(h, i1 - i0))
            if probe_every > 0 and not probe:
                for key, which in (("chebstep", 0), ("filterstep", 1)):
                    sec, ns, by = pv.kernel_probe(h, which)
                    probe[key] = dict(seconds=sec, samples=ns, bytes=by)
            t1 = time.perf_counter()
            if residual_check:
                r1, r2 = pv.residuals(h)
            else:
                k = pv.pevsl_get_nev(h)
                r1 = r2 = np.full(k, np.nan)
            t["residual"] += time.perf_counter() - t1
            lam, X = pv.pevsl_copy_result(h, mv.n, vectors)
            lams.append(lam)
            Xs.append(X)
            r1s.append(r1)
            r2s.append(r2)
        t["slices"] = time.perf_counter() - t0 - t["residual"]
    except BaseException:
        pv.pevsl_finish(h)
        raise
    if not keep:
        pv.pevsl_finish(h)
    lam = np.concatenate(lams)
    X = None if not vectors else np.concatenate(Xs, axis=0)
    r1 = np.concatenate(r1s)
    r2 = np.concatenate(r2s)
    order = np.argsort(lam, kind="stable")
    return EigResult(lam=lam[order], X=None if X is None else X[order], res_ref=r1[order], res2=r2[order],
                     lmin=lmin, lmax=lmax, xintv=xintv, deg=0, bar=float("nan"), lan_steps=steps,
                     filter_applies=applies, ncand=ncand, timings=t,
                     lan_steps_bounds=lb_steps, probe=probe, handle=h if keep else 0, slices=sli,
                     slice_counts=np.concatenate(counts), dos_count=ecnt, target_met=target_met,
                     waves=waves)


def solve_input(inputdir, basename, *, porder: int = 1, job: int = 1, lowfreq: float = 0.2, upfreq: float = 2.0,
                outputdir=None, bdeg="reference", tune=(1,), **kw):
    """The reference's main program on its own input files (src/mainnm.f90):
    read the mesh and model, assemble A and B (f1), setupmatvec,
    pnm_apply_pevsl on [lowfreq, upfreq] mHz, and, with ``outputdir``, write
    the mode files.  Returns (problem, setup, result, paths)."""
    from .assemble import assemble
    from .formats import read_input, write_modes
    M = read_input(inputdir, basename, porder)
    g0 = None
    if job == 2:
        if M.g0 is None:
            raise ValueError("JOB=2 needs the potential-acceleration file")
        g0 = M.g0 / 1.0e3                                  # m/s^2 -> km/s^2, src/mod_cg_create_matrix.f90:194
    P = assemble(M.mesh, rho=M.rho, vp=M.vp, vs=M.vs, g0=g0, job=job, name=basename)
    mv = setupmatvec(P, bdeg=bdeg, tune=tune)
    try:
        r = apply_pevsl(mv, lowfreq=lowfreq, upfreq=upfreq, **kw)
        paths = []
        if outputdir is not None and r.X is not None:
            paths = write_modes(outputdir, basename, job=job, porder=porder, lowfreq=lowfreq, upfreq=upfreq,
                                X_phys=r.physical_vectors(mv.d), vlist=M.mesh.vlist)
    finally:
        mv.free()
    return P, mv, r, paths
