#!/usr/bin/env python3
"""Benchmark of the filtered-Lanczos path on one MI355X.

Metric: seconds to all eigenpairs of A~ x = lambda B~ x in an interval, with
every pair at ||A~x - lambda B~x||_2/|lambda| <= 1e-10 (x B~-normalised).
A "step" is one complete solve starting from the device-resident operator:
Lanczos bounds of B~ and Chebyshev B-solve set-up, LanBounds of B~^-1 A~,
FindPol, ChebLanNr with Rayleigh-Ritz extraction and the residual check
(the span of src/mod_pevsl.f90:54-162 plus src/mod_matvec.f90:75-95).
Mesh generation, assembly and the host->HBM upload are outside the timed
region and reported separately.

Workload (default ``c2s``): the three-layer unit ball (rho 13->4, vp 11->8,
vs 3.5->6.5) with its outer core solid, seeded ~459k-tet isosurface-stuffed
P2 mesh, ~1.9M DoF; all eigenpairs in [189, 2845] (231 pairs).  The timed
step is B~ bounds and Chebyshev set-up, LanBounds, FindPol, block ChebLanNr
(4 columns as two 2-column filter chains on two streams) with extraction,
and the residual check; beside it the filter's cost per column and degree
by precision, the pieces of the same ball with its fluid outer core
(``c2``: set-up, LanBounds, FindPol, filter cost, product rates) and the
oracle's CPU baseline.  One solve takes minutes here, so the 2M-DoF
workloads time exactly one solve with no warm-up solve unless
``--long-solves 1``; the line records the steps asked for.  ``--workload
c1``: the homogeneous solid unit ball, seeded ~12k-tet P2 mesh, ~52k DoF,
the interval holding the ~30 lowest non-rigid modes.

Also reported on the c1 line:
  roofline      the dominant kernel (fused Chebyshev B-solve step), algorithmic
                bytes per launch over its mean duration from HIP events sampled
                on the solver's stream during the timed solves;
  spmv_hbm      sparse-product and Chebyshev-step rates at the 2M-DoF layered
                ball size (``--spmv-workload c2s``; ``none`` to skip);
  cpu_baseline  the CPU restatement (oracle, C + OpenMP) timed on a bounded
                sample of the same operator, scaled to a full solve;
  fp32_filter   one solve with FP32 B-solve values: time, per-eigenvalue
                difference from the FP64 solve, residuals.

``--kernel-only`` runs no solve: a copy kernel of known size, then the
dominant kernel alone (one B-solve launch per repetition), for hardware
counter runs (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE).  ``--spmv-only``
does the same for the sparse products at the ``--spmv-workload`` size.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
GATE = 1e-10


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# seconds to every eigenpair of the interval with all residuals <= 1e-10 (the
# FP64 solve; the FP32-filter solve and the SpMV HBM rates ride on the line)
METRIC = "s to all eigpairs in interval @resid\u22641e-10, FP32-filter vs FP64; SpMV HBM GB/s"


def traffic_from_profiles(kernel_tag):
    """HBM bytes per launch of the dominant kernel from the committed counter
    summary (a line 'traffic_bytes_per_launch: <kernel_tag> <bytes>')."""
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            parts = line.strip().strip("`").split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch:" and parts[1] == kernel_tag:
                try:
                    return float(parts[2]), f.name
                except ValueError:
                    pass
    return None, None


def upload_with_layouts(P, layouts, tune):
    """Upload A and B; autotune, or apply fixed product layouts.

    ``layouts``: 'auto', or comma-separated '<A|B><K>:<kind>:<params>' with
    kind 'lanes:<g>' (K = 1, g lanes per node row), 'rows:<r>' (r node rows
    per wavefront), 'tiled:<cap>:<r>[:<threads>]' (LDS tiles of at most cap
    nodes, workgroups of 1024 or 256 threads), 'groups:<g>' (K = 1, row
    groups, g per wavefront), 'wide:<g>' (K >= 2, scalar mass, g lanes per
    node row), 'twide:<cap>:<g>' (the same on LDS tiles) or 'pipe:<g>:<r>'
    (K = 1, 3x3 blocks, g lanes walking r rows)."""
    from normalmodes_amd import driver, pevsl as pv
    if layouts == "auto":
        return driver.upload(P, tune=tune)
    specs = parse_layouts(layouts)
    mv = driver.upload(P, tune=())
    for name, K, kind, p in specs:
        mat = mv.matA if name == "A" else mv.matB
        if kind == "lanes":
            pv.set_kernel_map(mat, group=p[0])
        elif kind == "rows":
            pv.set_kernel_map(mat, 0, K, p[0])
            if K == 1:
                pv.set_tiling(mat, 1, 0, 1)   # route single products through the row-block kernels
        elif kind == "tiled":
            pv.set_tiling(mat, K, p[0], p[1], p[2] if len(p) > 2 else 1024)
        elif kind == "groups":
            pv.set_row_groups(mat, p[0])
        elif kind == "wide":
            pv.set_wide(mat, K, p[0])
        elif kind == "pipe":
            pv.set_row_pipeline(mat, p[0], p[1])
        else:
            pv.set_tiled_wide(mat, K, p[0], p[1])
        mv.tuned[f"{name}{K}"] = {"fixed": ":".join([kind] + [str(v) for v in p])}
    return mv


_LAYOUT_ARGS = {"lanes": (1,), "rows": (1,), "tiled": (2, 3), "groups": (1,), "wide": (1,), "twide": (2,),
                "pipe": (2,)}


def parse_layouts(spec):
    """'<A|B><K>:<kind>:<params>,...' -> [(matrix, K, kind, [ints])]; raises
    ValueError on anything else, before any device work."""
    out = []
    for part in spec.split(","):
        f = part.strip().split(":")
        if len(f) < 2 or not f[0] or f[0][0] not in "AB" or f[1] not in _LAYOUT_ARGS:
            raise ValueError(f"bad layout {part!r}")
        K = int(f[0][1:] or 1)
        if K not in (1, 2, 4, 8, 16) or len(f) - 2 not in _LAYOUT_ARGS[f[1]]:
            raise ValueError(f"bad layout {part!r}")
        out.append((f[0][0], K, f[1], [int(v) for v in f[2:]]))
    return out


def kernel_only(args):
    """Copy kernel of known size, then the dominant kernel alone."""
    from normalmodes_amd import driver, pevsl as pv
    from normalmodes_amd.workloads import build_problem
    P = build_problem(args.workload)
    kpart = args.block // args.split if args.block > 1 and args.split > 1 else args.block
    if isinstance(P, driver.FluidProblem):
        mv = driver.setupmatvec_fs(P, bdeg="auto" if args.bdeg == "auto" else args.bdeg, apdeg=args.apdeg,
                                   tune=sorted({1, kpart}))
    else:
        mv = upload_with_layouts(P, args.layouts, sorted({1, kpart}))
        driver.setup_bsolve(mv, bdeg="auto" if args.bdeg == "auto" else args.bdeg)
    n_copy = 1 << 24   # 16M doubles: 128 MiB read + 128 MiB written per launch
    t_copy = pv.bench_copy(n_copy, 5)
    info = pv.chebiter_info(mv.chebB)
    if (args.refine or P.n > 500000) and not isinstance(P, driver.FluidProblem):
        # the headline's filter application itself (B-solves, refined or not,
        # on args.split concurrent chains): its Chebyshev steps are the
        # dominant kernel, here in the same concurrent arrangement as in the
        # timed solve
        from normalmodes_amd.workloads import INTERVALS
        degs = driver.refine_bsolves(mv, passes=args.refine)
        h = driver._open_solver(mv, precision=pv.FP64, seed=7, rtol=1e-9, block=args.block, split=args.split)
        try:
            lmin, lmax = pv.pevsl_lanbounds(h, 3000, 5000, 1e-5)
            a, b = INTERVALS[args.workload]
            pol = pv.pevsl_findpol(np.array([a, b, lmin, lmax]), 0.8, 0.7)
            deg = pv.pol_info(pol).deg
            t = pv.bench_filter(h, pol, args.block, reps=args.steps)
            pv.pevsl_freepol(pol)
        finally:
            pv.pevsl_finish(h)
        print(json.dumps({"copy": {"doubles": n_copy, "bytes_per_launch": 16 * n_copy, "launches": 6,
                                   "us": round(t_copy * 1e6, 2)},
                          "filter": {"workload": args.workload, "n_dof": P.n, "block": args.block,
                                     "chains": args.split, "filter_degree": deg, "bdeg": mv.bdeg,
                                     "refine_inner_degree": degs, "applications": args.steps + 1,
                                     "seconds_per_application": round(t, 4)}}), flush=True)
        mv.free()
        return
    if kpart > 1:   # the k-column Chebyshev step of one part, repeated
        reps = args.steps * mv.bdeg
        t, b = pv.bench_block_chebstep(mv.chebB, kpart, reps)
        kern = {"kernel": f"k-column Chebyshev step, K={kpart}", "launches": reps + 1}
    else:
        t, b = pv.bench_bsolve(mv.chebB, args.steps)
        kern = {"kernel": "B-solve (deg single-column steps)", "launches": args.steps + 1}
    print(json.dumps({"copy": {"doubles": n_copy, "bytes_per_launch": 16 * n_copy, "launches": 6,
                               "us": round(t_copy * 1e6, 2)},
                      "bsolve": {"workload": args.workload, "n_dof": P.n, "deg": mv.bdeg, "block": args.block,
                                 "us": round(t * 1e6, 2), "algorithmic_bytes": b, **kern, **info}}), flush=True)
    mv.free()


def solve_once(mv, interval, bdeg, precision, probe_every, block, split=0, refine=0, verbose=0):
    from normalmodes_amd import driver, pevsl as pv
    t0 = time.perf_counter()
    driver.setup_bsolve(mv, bdeg=bdeg)
    driver.refine_bsolves(mv, passes=refine)
    r = driver.apply_pevsl(mv, interval=interval, precision=precision, vectors=False,
                           probe_every=probe_every, block=block, split=split if block > 1 else 0,
                           verbose=verbose)
    return time.perf_counter() - t0, r


FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (matrix and vector)


def reorth_rate(n, m, k, reps=10):
    """The block re-orthogonalisation products on the FP64 matrix cores
    (hand-written v_mfma_f64_16x16x4f64 kernels): H = V^T W and W -= Z H
    (V, Z n x m, W n x k).  Rates against the FP64 peak (MFMA utilisation)
    and against 8 TB/s, with the rocBLAS / FMA kernels they replace."""
    from normalmodes_amd import pevsl as pv
    out = dict(n=n, m=m, k=k)
    for op, name in ((1, "vtw"), (0, "w_minus_zh")):
        for kernel, tag in ((1, "mfma"), (0, "library")):
            s, b, f = pv.bench_dense(op, kernel, n, m, k, reps)
            out[f"{name}_{tag}"] = dict(us=round(s * 1e6, 2), tflops=round(f / s / 1e12, 2),
                                        frac_fp64_peak=round(f / s / 1e12 / FP64_PEAK_TFLOPS, 4),
                                        gbs=round(b / s / 1e9, 1), frac_hbm=round(b / s / 1e9 / HBM_PEAK_GBS, 3))
    return out


def spmv_rates(name, reps=50):
    from normalmodes_amd import driver, pevsl as pv
    from normalmodes_amd.workloads import build_problem
    t0 = time.perf_counter()
    P = build_problem(name)
    t_build = time.perf_counter() - t0
    mv = driver.upload(P, tune=(1,))
    driver.setup_bsolve(mv, bdeg="auto")
    out = dict(workload=name, n_dof=P.n, nnz_A=P.A.nnz, nnz_B=P.B.nnz, build_s=round(t_build, 2),
               kernel_layouts=mv.tuned)
    try:
        for key, mat in (("A_bsr3", mv.matA), ("B_mass3", mv.matB)):
            for fp32 in (False, True):
                s, b = pv.bench_spmv(mat, reps, fp32)
                out[f"{key}{'_fp32' if fp32 else ''}"] = dict(us=round(s * 1e6, 2), gbs=round(b / s / 1e9, 1),
                                                               frac=round(b / s / 1e9 / HBM_PEAK_GBS, 3))
        for fp32 in (False, True):
            s, b = pv.bench_chebstep(mv.chebB, reps, fp32)
            out[f"chebstep{'_fp32' if fp32 else ''}"] = dict(us=round(s * 1e6, 2), gbs=round(b / s / 1e9, 1),
                                                             frac=round(b / s / 1e9 / HBM_PEAK_GBS, 3))
    finally:
        mv.free()
    out["reorth_gemm"] = reorth_rate(P.n, 400, 8)
    out["reorth_gemm_k64"] = reorth_rate(P.n, 512, 64, reps=5)
    return out


WORKLOAD_TEXT = {
    "c1": ("c1: homogeneous solid unit ball (rho=1, vp=1.73, vs=1), seeded jittered-lattice Delaunay mesh of "
           "{tets} P2 tets, {n} DoF, all eigenpairs in [{a:.6g}, {b:.6g}]"),
    "tiny": "tiny: homogeneous solid unit ball, {tets} P2 tets, {n} DoF, all eigenpairs in [{a:.6g}, {b:.6g}]",
    "c2s": ("c2s: the 2M-DoF three-layer unit ball (rho 13->4, vp 11->8, vs 3.5->6.5 piecewise linear in r, "
            "interfaces at r = 0.19 and 0.55 as tetrahedron faces) with the outer core solid, seeded "
            "isosurface-stuffed BCC mesh of {tets} P2 tets, {n} DoF, all eigenpairs in [{a:.6g}, {b:.6g}] "
            "(231 pairs, the 5-fold 218.4 multiplet up to 2830.2)"),
}


# the 2M-DoF three-layer ball with a fluid outer core (JOB=2, Earth-sized,
# radius 6371 km): the band of its ~200 seismic modes in mHz, placed clear of
# the fluid's cluster of ~9e4 near-zero (undertone) eigenvalues below 0.13 mHz
C2_BAND_MHZ = (0.8, 1.3)


def product_rates(mv, reps=20):
    """Rates of every product and Chebyshev step of a fluid-solid set-up."""
    from normalmodes_amd import pevsl as pv
    rates = {}
    for key, mat in (("Ad", mv.fs.matAd), ("B", mv.matB), ("Ap", mv.fs.matAp), ("E", mv.fs.matE),
                     ("ET", mv.fs.matET)):
        for fp32 in (False, True):
            sec, b = pv.bench_spmv(mat, reps, fp32)
            rates[key + ("_fp32" if fp32 else "")] = dict(layout=pv.csr_info(mat)["layout"],
                                                          us=round(sec * 1e6, 2), gbs=round(b / sec / 1e9, 1),
                                                          frac=round(b / sec / 1e9 / HBM_PEAK_GBS, 3))
    for key, cheb in (("B_chebstep", mv.chebB), ("Ap_chebstep", mv.fs.chebAp)):
        for fp32 in (False, True):
            sec, b = pv.bench_chebstep(cheb, reps, fp32)
            rates[key + ("_fp32" if fp32 else "")] = dict(us=round(sec * 1e6, 2), gbs=round(b / sec / 1e9, 1),
                                                          frac=round(b / sec / 1e9 / HBM_PEAK_GBS, 3))
    return rates


def filter_rates(mv, interval, block, split, lanbounds_tol, reps=1):
    """Seconds of one filter application per column and degree, by precision
    mode, at the solve's block width and stream split and at one column."""
    from normalmodes_amd import driver, pevsl as pv
    out = {}
    h = driver._open_solver(mv, precision=pv.FP64, seed=7, rtol=1e-9)
    try:
        lmin, lmax = pv.pevsl_lanbounds(h, 3000, 5000, lanbounds_tol)
        pol = pv.pevsl_findpol(np.array([interval[0], interval[1], lmin, lmax]), 0.8, 0.7)
        deg = pv.pol_info(pol).deg
        try:
            for label, prec, passes in (("fp64", pv.FP64, 0), ("fp64_refined", pv.FP64, 2),
                                        ("fp32_values", pv.FP32_FILTER, 0), ("fp32_filter_A", pv.FP32_FILTER_A, 0),
                                        ("fp32_iterates", pv.FP32_BSOLVE, 0)):
                pv.pevsl_set_precision(h, prec)
                driver.refine_bsolves(mv, passes=passes)
                row = {}
                for K, sp in ((1, 0), (block, split)):
                    pv.set_split_streams(h, sp)
                    sec = pv.bench_filter(h, pol, K, reps=reps)
                    row[f"K{K}"] = round(sec / K / deg * 1e3, 4)
                out[label] = row
        finally:
            driver.refine_bsolves(mv, passes=0)
            pv.pevsl_freepol(pol)
    finally:
        pv.pevsl_finish(h)
    out["unit"] = "ms per column and filter degree"
    out["degree"] = deg
    return out


def run_fluid(args):
    """The 2M-DoF fluid-solid headline: setupmatvec_fs, then timed solves of
    the band (B~ bounds and Chebyshev set-up, LanBounds, FindPol, ChebLanNr
    with extraction, residual check), FP64 for ``value`` and one FP32-filter
    solve beside it."""
    from normalmodes_amd import driver, pevsl as pv
    from normalmodes_amd.workloads import build_problem
    t0 = time.perf_counter()
    Q = build_problem(args.workload)
    t_build = time.perf_counter() - t0
    band = tuple(float(v) for v in args.band_mhz.split(","))
    interval = (driver.freq_to_lambda(band[0]), driver.freq_to_lambda(band[1]))
    block = args.block or 4
    split = 2 if args.split is None else args.split
    kpart = block // split if split > 1 and block > 1 else block
    log(f"[bench] built {args.workload}: n={Q.n} pressures={Q.n_p} tets={Q.mesh.n_elem} in {t_build:.1f}s")
    t0 = time.perf_counter()
    mv = driver.setupmatvec_fs(Q, bdeg="auto", apdeg=args.apdeg, tune=sorted({1, kpart, block}), verbose=True)
    t_setup = time.perf_counter() - t0
    lb = (3000, 5000, args.lanbounds_tol)

    def solve(precision, probe_every, refine):
        t1 = time.perf_counter()
        driver.setup_bsolve(mv, bdeg="auto")
        driver.refine_bsolves(mv, passes=refine)
        r = driver.apply_pevsl(mv, interval=interval, precision=precision, vectors=False, block=block,
                               split=split, probe_every=probe_every, lanbounds=lb, verbose=max(1, 24 // block))
        return time.perf_counter() - t1, r

    for i in range(args.warmup):
        el, r = solve(pv.FP64, 0, args.refine)
        log(f"[bench] warmup {i}: {el:.1f}s nev={r.lam.size}")
    times, r = [], None
    gate_all = True
    for i in range(args.steps):
        el, r = solve(pv.FP64, args.probe_every, args.refine)
        times.append(el)
        gate_all = gate_all and bool(r.res2.size > 0 and np.all(r.res2 <= GATE))
        log(f"[bench] step {i}: {el:.1f}s nev={r.lam.size} deg={r.deg} columns={r.lan_steps} "
            f"max res2={np.max(r.res2) if r.res2.size else float('nan'):.2e}")
    T = max(times)
    gate_ok = gate_all
    chk = r.probe.get("chebstep", {"seconds": 0.0, "bytes": 0.0, "samples": 0})
    achieved = chk["bytes"] / chk["seconds"] / 1e9 if chk["seconds"] > 0 else 0.0
    kname = f"block{kpart}_chebstep" if kpart > 1 else "mass3_chebstep"
    traffic, traffic_src = traffic_from_profiles(f"{kname}:{args.workload}")
    line = {
        "metric": METRIC, "value": round(T, 3), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(T * 1e3, 1), "higher_is_better": False, "scaling": "weak",
        "steps_requested": args.requested,
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": (f"{args.workload}: three-layer ball of radius 6371 km (solid inner core r<0.19, fluid "
                         f"outer core 0.19-0.55 with vs=0 and pressure DoFs, solid mantle; rho 13->4, vp 11->8, "
                         f"vs 3.5->6.5 piecewise linear in r), JOB=2 with its own reference gravity, seeded "
                         f"isosurface-stuffed BCC mesh of {Q.mesh.n_elem} P2 tets, {Q.n} displacement DoF + "
                         f"{Q.n_p} pressures, all eigenpairs in [{band[0]}, {band[1]}] mHz"),
            "n_dof": Q.n, "n_pressure": Q.n_p, "n_tets": Q.mesh.n_elem,
            "nnz": {k: getattr(Q, k).nnz for k in ("Ad", "B", "E", "ET", "Ap")},
            "band_mhz": list(band), "interval": list(interval), "nev": int(r.lam.size),
            "filter_degree": r.deg, "bsolve_degree": mv.bdeg, "ap_degree": mv.fs.apdeg,
            "lanczos_columns": r.lan_steps, "lanbounds": [r.lmin, r.lmax], "lanbounds_tol": args.lanbounds_tol,
            "block": block, "filter_chains": split if split > 1 else 1, "refine_passes": args.refine,
            "kernel_layouts": mv.tuned, "parallelism": "1 GPU", "global_batch": 1, "seq_len": Q.n,
        },
        "gate": {"resid_max": float(np.max(r.res2)) if r.res2.size else None,
                 "resid_ref_max": float(np.max(r.res_ref)) if r.res_ref.size else None,
                 "passed": gate_ok, "threshold": GATE, "target_met": bool(r.target_met)},
        "phases_s": {k: round(v, 3) for k, v in r.timings.items()},
        "setup_s": {"build": round(t_build, 2), "setupmatvec_fs": round(t_setup, 2)},
        "roofline": {
            "bound": "hbm",
            "kernel": (f"fused {kpart}-column Chebyshev B-solve step (scalar mass x I3 product + update on "
                       f"{kpart} interleaved columns) of one filter chain; {split if split > 1 else 1} chains "
                       f"run at once on as many streams"),
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
            "bytes_per_launch": chk["bytes"], "mean_launch_us": round(chk["seconds"] * 1e6, 3),
            "samples": chk["samples"],
            "concurrent_parts": split if split > 1 else 1,
            "aggregate_achieved": round(achieved * (split if split > 1 else 1), 1),
        },
    }
    if args.fp32_compare:
        try:
            mode = {"values": pv.FP32_FILTER, "filter_A": pv.FP32_FILTER_A, "iterates": pv.FP32_BSOLVE}[args.fp32_mode]
            el32, r32 = solve(mode, 0, 0)
            m = min(r32.lam.size, r.lam.size)
            line["fp32_filter"] = {
                "seconds": round(el32, 3), "mode": args.fp32_mode, "nev": int(r32.lam.size),
                "speedup_vs_fp64": round(T / el32, 3),
                "max_rel_eig_diff_vs_fp64": float(np.max(np.abs(r32.lam[:m] - r.lam[:m]) / np.abs(r.lam[:m])))
                if m and r32.lam.size == r.lam.size else None,
                "resid_max": float(np.max(r32.res2)) if r32.res2.size else None,
                "lanczos_columns": r32.lan_steps}
            log(f"[bench] fp32 filter: {line['fp32_filter']}")
        except Exception as e:
            line["fp32_filter"] = {"error": str(e)}
    try:
        line["filter_cost"] = filter_rates(mv, interval, block, split, args.lanbounds_tol)
        log(f"[bench] filter cost: {line['filter_cost']}")
    except Exception as e:
        line["filter_cost"] = {"error": str(e)}
    try:
        line["spmv_hbm"] = product_rates(mv)
        log(f"[bench] product rates: {line['spmv_hbm']}")
    except Exception as e:
        line["spmv_hbm"] = {"error": str(e)}
    try:
        line["reorth_gemm"] = reorth_rate(Q.n, min(r.lan_steps, 512), block)
    except Exception as e:
        line["reorth_gemm"] = {"error": str(e)}
    if args.cpu_seconds > 0:
        try:
            line["cpu_baseline"] = cpu_baseline_fluid(Q, mv, r, args.cpu_seconds)
            log(f"[bench] cpu baseline: {line['cpu_baseline']}")
        except Exception as e:
            line["cpu_baseline"] = {"error": str(e)}
    mv.free()
    print(json.dumps(line), flush=True)


def producer_compare(Q, device=True):
    """The fluid-solid producer on the GPU (nmg_assemble_fs) against the CPU
    one on the same body and materials: seconds of each (host pattern work
    included in both) and the largest difference relative to each block's
    largest entry.  ``device=False`` runs the CPU producer on both sides."""
    from normalmodes_amd.assemble import assemble_fs
    vp = np.sqrt((Q.lam + 2 * Q.mu) / Q.rho)
    vs = np.sqrt(Q.mu / Q.rho)
    kw = dict(rho=Q.rho, vp=vp, vs=vs, g0=Q.g0, job=Q.job)
    t0 = time.perf_counter()
    D = assemble_fs(Q.mesh, device=device, **kw)
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    C = assemble_fs(Q.mesh, **kw)
    t_cpu = time.perf_counter() - t0
    same = all(np.array_equal(getattr(D, k).col, getattr(C, k).col) and
               np.array_equal(getattr(D, k).rowptr, getattr(C, k).rowptr) for k in ("Ad", "B", "E", "ET", "Ap"))
    err = max(float(np.abs(getattr(D, k).val - getattr(C, k).val).max() / np.abs(getattr(C, k).val).max())
              for k in ("Ad", "B", "E", "ET", "Ap"))
    return dict(gpu_s=round(t_dev, 3), cpu_s=round(t_cpu, 3),
                cpu_threads=int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())), pattern_same=same,
                max_rel_diff=err)


def fluid_pieces(args):
    """Measured pieces of the fluid-solid ball (c2): assembly and set-up,
    LanBounds, the FindPol filter the reference's rule picks for the band,
    the filter's cost per column and degree by precision, and every product
    rate.  Its whole solve is not run: FindPol's end-of-spectrum rule makes
    the filter of any band of the lowest seismic modes a low-pass one over
    [lmin, b], which passes the fluid's ~1e5 near-zero eigenvalues."""
    from normalmodes_amd import driver, pevsl as pv
    from normalmodes_amd.workloads import build_problem
    t0 = time.perf_counter()
    Q = build_problem("c2")
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    mv = driver.setupmatvec_fs(Q, bdeg="auto", apdeg="auto", tune=(1, 2), verbose=True)
    t_setup = time.perf_counter() - t0
    band = C2_BAND_MHZ
    interval = (driver.freq_to_lambda(band[0]), driver.freq_to_lambda(band[1]))
    out = dict(workload=("c2: the three-layer ball of radius 6371 km with its fluid outer core (vs = 0, pressure "
                         "DoFs), JOB=2 with its reference gravity, stuffed mesh"),
               n_dof=Q.n, n_pressure=Q.n_p, n_tets=Q.mesh.n_elem,
               nnz={k: getattr(Q, k).nnz for k in ("Ad", "B", "E", "ET", "Ap")},
               build_s=round(t_build, 2), setup_s=round(t_setup, 2), bsolve_degree=mv.bdeg, ap_degree=mv.fs.apdeg,
               band_mhz=list(band))
    try:
        out["producer"] = producer_compare(Q)
        log(f"[bench] c2 producer: {out['producer']}")
    except Exception as e:
        out["producer"] = {"error": str(e)}
    try:
        h = driver._open_solver(mv, precision=pv.FP64, seed=7, rtol=1e-9)
        try:
            t0 = time.perf_counter()
            lmin, lmax = pv.pevsl_lanbounds(h, 3000, 5000, 1e-5)
            out["lanbounds"] = dict(seconds=round(time.perf_counter() - t0, 2), lmin=lmin, lmax=lmax,
                                    steps=pv.solve_stats(h)["lan_steps"])
            pol = pv.pevsl_findpol(np.array([interval[0], interval[1], lmin, lmax]), 0.8, 0.7)
            info = pv.pol_info(pol)
            ftype = 1 if info.gam <= -1.0 + 1e-12 else (2 if info.gam >= 1.0 - 1e-12 else 0)
            out["findpol"] = dict(degree=info.deg, type=ftype, gamma=info.gam,
                                  note="type 1: the interval touches the lower end of the spectrum (within "
                                       "5e-4 of its width), so the filter is low-pass over [lmin, b]")
            pv.pevsl_freepol(pol)
        finally:
            pv.pevsl_finish(h)
        out["filter_cost"] = filter_rates(mv, interval, 4, 2, 1e-5)
        out["rates"] = product_rates(mv, reps=10)
    finally:
        mv.free()
    return out


def fluid_rates_only(args):
    """``--spmv-only`` for a fluid-solid workload (c2, c4): the operator's
    product and Chebyshev-step rates, the k-column B~ step at the slice widths
    16, 32 and 64, one filter application on 16 columns as four chains, and
    the slice-wave plan of the Lanczos bases for 16 slices of 20-250 pairs."""
    from normalmodes_amd import driver, pevsl as pv
    from normalmodes_amd.workloads import build_problem
    t0 = time.perf_counter()
    Q = build_problem(args.workload)
    t_build = time.perf_counter() - t0
    log(f"[bench] built {args.workload}: n={Q.n} pressures={Q.n_p} in {t_build:.1f}s")
    t0 = time.perf_counter()
    mv = driver.setupmatvec_fs(Q, bdeg="auto", apdeg="auto", tune=(1, 4, 16), verbose=True)
    t_setup = time.perf_counter() - t0
    log(f"[bench] setupmatvec_fs {t_setup:.1f}s")
    mats = {k: getattr(Q, k) for k in ("Ad", "B", "E", "ET", "Ap")}
    out = dict(workload=args.workload, n_dof=Q.n, n_pressure=Q.n_p, n_tets=Q.mesh.n_elem,
               nnz={k: M.nnz for k, M in mats.items()},
               matrix_gbytes_fp64_csr=round(sum(M.nnz * 12 + (M.n_rows + 1) * 8 for M in mats.values()) / 1e9, 2),
               build_s=round(t_build, 1), setup_s=round(t_setup, 1), bsolve_degree=mv.bdeg, ap_degree=mv.fs.apdeg,
               rates=product_rates(mv, reps=max(3, args.steps)))
    log(f"[bench] rates {out['rates']}")
    blk = {}
    for K in (16, 32, 64):
        for mode, tag in ((0, "fp64"), (1, "fp32_values"), (2, "fp32_iterates")):
            sec, b = pv.bench_block_chebstep(mv.chebB, K, reps=max(3, args.steps), fp32=mode)
            blk[f"K{K}_{tag}"] = dict(us=round(sec * 1e6, 1), us_per_column=round(sec * 1e6 / K, 2),
                                      gbs=round(b / sec / 1e9, 1), frac=round(b / sec / 1e9 / HBM_PEAK_GBS, 3))
        log(f"[bench] block chebstep K={K}: {blk}")
    out["block_chebstep"] = blk
    h = driver._open_solver(mv, precision=pv.FP64, seed=7, rtol=1e-9, block=16, split=4)
    try:
        lmin, lmax = pv.pevsl_lanbounds(h, 3000, 5000, 1e-5)
        band = C2_BAND_MHZ
        pol = pv.pevsl_findpol(np.array([driver.freq_to_lambda(band[0]), driver.freq_to_lambda(band[1]), lmin,
                                         lmax]), 0.8, 0.7)
        deg = pv.pol_info(pol).deg
        sec = pv.bench_filter(h, pol, 16, reps=1)
        out["filter_16_columns"] = dict(seconds=round(sec, 2), degree=deg, chains=4,
                                        ms_per_column_degree=round(sec / 16 / deg * 1e3, 3), lanbounds=[lmin, lmax])
        pv.pevsl_freepol(pol)
    finally:
        pv.pevsl_finish(h)
    counts = [20 + (230 * i) // 15 for i in range(16)]    # 16 slices of 20..250 pairs
    out["slice_waves_200GB"] = driver.plan_slice_waves(counts, Q.n, 200e9)
    out["basis_gbytes_all_slices"] = round(sum(3.0 * c + 40 for c in counts) * 16.0 * Q.n / 1e9, 1)
    mv.free()
    return out


def cpu_baseline_fluid(Q, mv, r, seconds_budget):
    """The oracle's fluid-solid pencil (CPU restatement of sparsefsAV and the
    filter, C + OpenMP primitives) timed on a bounded sample: degrees of the
    same filter on the same operator; scaled by the filter degrees of the
    GPU solve, counted per column (a single-vector CPU solve applies the
    filter once per Lanczos vector)."""
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc
    F = orc.FluidPencil(Q, bdeg=mv.bdeg, bounds_B=(mv.blmin, mv.blmax), apdeg=mv.fs.apdeg,
                        bounds_Ap=(mv.fs.aplmin, mv.fs.aplmax))
    pol = orc.findpol(r.xintv, 0.8, 0.7)
    z = orc.start_vector(Q.n, 7)
    small = orc.Pol(**{**pol.__dict__, "deg": 1, "mu": pol.mu[:2].copy()})
    t0 = orc.wtime()
    F.chebav(small, z)
    per_deg = orc.wtime() - t0
    m = int(max(1, min(pol.deg, seconds_budget / max(per_deg, 1e-6))))
    samp = orc.Pol(**{**pol.__dict__, "deg": m, "mu": pol.mu[:m + 1].copy()})
    t0 = orc.wtime()
    F.chebav(samp, z)
    t_s = orc.wtime() - t0
    per_deg = t_s / m
    degrees = r.deg * r.filter_applies + r.lan_steps_bounds
    return dict(value=round(per_deg * degrees, 1), unit="s", cores=orc.num_threads(), kind="port",
                cpu_model=cpu_model(), host_threads=os.cpu_count(),
                sample=(f"oracle fluid-solid filter of degree {m} (of {pol.deg}; B-solve degree {mv.bdeg}, "
                        f"Ap solve degree {mv.fs.apdeg}) on the same {Q.n}-DoF operator: {t_s:.1f} s; scaled by "
                        f"the {degrees} filter degrees of the GPU solve ({r.filter_applies} filtered columns x "
                        f"{r.deg} + {r.lan_steps_bounds} LanBounds steps)"))


def cpu_baseline(P, mv, r, seconds_budget):
    """Time the oracle (CPU restatement) on a bounded sample of the same
    operator: degrees of the same filter, each one A~ product plus one
    Chebyshev B-solve of the same degree; scale to the full solve by the
    filter degrees and Lanczos steps the GPU run took."""
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc
    threads = orc.num_threads()
    Pn = orc.Pencil(P.A, P.B, bdeg=mv.bdeg, bounds_B=(mv.blmin, mv.blmax))
    pol = orc.findpol(r.xintv, 0.8, 0.7)
    # calibrate on a 2-degree filter, then size the sample to the budget
    z = orc.start_vector(P.n, 7)
    small = orc.Pol(**{**pol.__dict__, "deg": 2, "mu": pol.mu[:3].copy()})
    t0 = orc.wtime()
    Pn.chebav(small, z)
    per_deg = (orc.wtime() - t0) / 2.0
    m = int(max(2, min(pol.deg, seconds_budget / max(per_deg, 1e-6))))
    samp = orc.Pol(**{**pol.__dict__, "deg": m, "mu": pol.mu[:m + 1].copy()})
    t0 = orc.wtime()
    Pn.chebav(samp, z)
    t_s = orc.wtime() - t0
    per_deg = t_s / m
    # full solve: filter degrees of ChebLanNr + one product pair per LanBounds
    # step (already counted in the GPU's filter_applies) -> per-degree time
    degrees = r.deg * r.filter_applies + r.lan_steps_bounds
    est = per_deg * degrees
    # CPU SpMV rates of the same matrices (oracle CSR product, all threads)
    x = np.random.default_rng(1).standard_normal(P.n)
    rates = {}
    for key, M in (("A", P.A), ("B", P.B)):
        orc.spmv(M, x)
        reps = 20
        t0 = orc.wtime()
        for _ in range(reps):
            orc.spmv(M, x)
        sec = (orc.wtime() - t0) / reps
        byt = M.nnz * 12.0 + (M.n_rows + 1) * 8.0 + 2 * 8.0 * M.n_rows
        rates[key] = dict(us=round(sec * 1e6, 1), gbs=round(byt / sec / 1e9, 2))
    return dict(value=round(est, 2), unit="s", cores=threads, kind="port", cpu_model=cpu_model(),
                host_threads=os.cpu_count(), spmv_csr_fp64=rates,
                sample=(f"oracle filter of degree {m} (of {pol.deg}) with B-solve degree {mv.bdeg} on the "
                        f"same {P.n}-DoF operator: {t_s:.1f} s; scaled by the {degrees} filter degrees "
                        f"(ChebLanNr {r.filter_applies} applications x {r.deg} + {r.lan_steps_bounds} "
                        f"LanBounds steps) of the GPU solve"))


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_full_solve(P, interval, bdeg):
    """One complete single-vector solve of the CPU restatement (oracle) on the
    same operator and interval: the measurement the bench's scaled estimate
    is checked against."""
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc
    orc.set_verbose(20)
    t0 = orc.wtime()
    lam, _, info = orc.solve_interval(P.A, P.B, interval, bdeg=bdeg, return_info=True)
    sec = orc.wtime() - t0
    orc.set_verbose(0)
    return dict(seconds=round(sec, 2), nev=int(lam.size), cores=orc.num_threads(), cpu_model=cpu_model(),
                info={k: (v if isinstance(v, (int, float)) else str(v)) for k, v in info.items()
                      if k in ("nsteps", "deg", "lmin", "lmax", "lanbounds_steps", "ncand")})


def resolve_defaults(args):
    """Per-workload defaults of the flags left unset; returns whether the
    workload is a fluid-solid body.  c2s (the headline) runs block 4 as two
    2-column chains, plain FP64, no warm-up solve, and measures the fluid
    ball's pieces beside it; c1 keeps its block 8 and its comparisons."""
    fluid = args.workload in ("c2", "c2r", "c4")
    big = args.workload == "c2s"
    if args.warmup is None:
        args.warmup = 0 if fluid or big else 1
    args.requested = {"steps": args.steps, "warmup": args.warmup}
    args.capped = False
    if (big or fluid) and not (args.kernel_only or args.spmv_only or args.long_solves):
        # one whole solve at 2M DoF takes about nine minutes: time exactly one,
        # with no warm-up solve (the set-up's autotuner has already launched
        # the product and B-solve kernels), whatever --steps / --warmup ask for
        if args.steps != 1 or args.warmup != 0:
            log(f"[bench] {args.workload}: a whole solve takes minutes, so this run times 1 solve with no "
                f"warm-up solve (asked: --steps {args.steps} --warmup {args.warmup}; --long-solves 1 keeps them)")
            args.capped = True
        args.steps, args.warmup = 1, 0
    if not fluid:
        args.block = args.block or (4 if big else 8)
        args.split = 2 if args.split is None else args.split
    args.refine = 0 if args.refine is None else args.refine
    for name, small_default in (("single_compare", 1), ("fp32_compare", 1), ("adjoint", 1)):
        if getattr(args, name) is None:
            setattr(args, name, 0 if (big or fluid) else small_default)
    if args.fluid_pieces is None:
        args.fluid_pieces = 1 if big else 0
    if args.spmv_workload is None:
        args.spmv_workload = "none" if big or fluid else "c2s"
    return fluid


def build_parser():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed full solves")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed full solves before timing (default 0 for c2, 1 otherwise)")
    ap.add_argument("--workload", default="c2s", choices=["c2s", "c2", "c2r", "c4", "c1", "tiny"],
                    help="c2s: the 2M-DoF three-layer ball (headline whole solve); c2: the same ball with its "
                         "fluid outer core (measured pieces; its whole solve is beyond reach, see DESIGN); "
                         "c1: the 52k-DoF solid ball")
    ap.add_argument("--band-mhz", default=",".join(str(v) for v in C2_BAND_MHZ),
                    help="band of the fluid-solid workloads, 'lo,hi' in mHz")
    ap.add_argument("--apdeg", default="auto", help="Ap~ Chebyshev degree of the fluid-solid set-up: auto | reference")
    ap.add_argument("--refine", type=int, default=None,
                    help="mixed-precision refinement passes of the Chebyshev solves in the timed solve (default 0: "
                         "plain FP64; 2 passes floor the residual near 4e-10 at 2M DoF, above the gate)")
    ap.add_argument("--fp64-compare", type=int, default=0,
                    help="also time one plain FP64 solve beside a refined headline (several minutes at c2s)")
    ap.add_argument("--fluid-pieces", type=int, default=None,
                    help="with c2s, also measure the fluid-solid ball's set-up, bounds, filter and product costs "
                         "(default 1 for c2s)")
    ap.add_argument("--lanbounds-tol", type=float, default=1e-5, help="LanBounds tolerance (the reference's 1e-5)")
    ap.add_argument("--fp32-mode", default="filter_A", choices=["values", "filter_A", "iterates"],
                    help="precision of the FP32-filter solve: B~ values; + stiffness values; B~ values + iterates")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "fp32a", "fp32i"],
                    help="timed solve: fp64; fp32 = FP32 B~ values in the B-solve; fp32a = and FP32 stiffness "
                         "values in the filter; fp32i = FP32 B~ values and iterates in the k-column B-solves")
    ap.add_argument("--bdeg", default="auto", help="B-solve degree: auto | reference | <int>")
    ap.add_argument("--probe-every", type=int, default=64)
    ap.add_argument("--block", type=int, default=0,
                    help="Lanczos block width (1 = the reference's single-vector ChebLanNr; the block variant "
                         "reorthogonalises with FP64 GEMMs and streams the matrices once per K columns)")
    ap.add_argument("--split", type=int, default=None,
                    help="run each k-column B-solve as this many concurrent parts on as many streams (2 or 4; "
                         "0 = one stream)")
    ap.add_argument("--single-compare", type=int, default=None,
                    help="with --block > 1, also time one single-vector (block 1) solve (0 skips)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline budget; 0 skips")
    ap.add_argument("--cpu-full", type=int, default=0,
                    help="also time one complete single-vector solve of the CPU restatement (several minutes)")
    ap.add_argument("--spmv-workload", default=None,
                    help="size for the SpMV HBM rates; none skips (default: the headline's own matrices for c2s, "
                         "c2s for c1)")
    ap.add_argument("--fp32-compare", type=int, default=None, help="also run one FP32-filter solve (0 skips)")
    ap.add_argument("--long-solves", type=int, default=0,
                    help="1: run --steps / --warmup whole solves as asked on the 2M-DoF workloads (each takes "
                         "minutes); 0 (default): time exactly one solve there, with no warm-up solve")
    ap.add_argument("--kernel-only", action="store_true", help="copy kernel + dominant kernel only (counter runs)")
    ap.add_argument("--spmv-only", action="store_true",
                    help="copy kernel + the --spmv-workload product rates only, --steps launches each (counter runs)")
    ap.add_argument("--adjoint", type=int, default=None, help="time the adjoint of the solved pairs (0 skips)")
    ap.add_argument("--layouts", default="auto",
                    help="'auto': autotune the product layouts on the device; or fixed lanes per node row as "
                         "'A:<g>,B:<g>' (g in 8,16,32,64), e.g. for profiler runs, where timing-based tuning "
                         "is perturbed")
    return ap


def main(argv=None):
    args = build_parser().parse_args(argv)

    if args.gpus != 1:
        log("nmgpu runs on one GPU: collectives and sharding are not part of it "
            "(the path would shard by spectrum slices, one group per GPU); run with --gpus 1")
        sys.exit(2)
    bdeg = args.bdeg if args.bdeg in ("auto", "reference") else int(args.bdeg)
    fluid = resolve_defaults(args)
    if fluid and not (args.kernel_only or args.spmv_only):
        run_fluid(args)
        return
    if args.kernel_only:
        args.bdeg = bdeg
        kernel_only(args)
        return
    if args.spmv_only and fluid:
        print(json.dumps(fluid_rates_only(args)), flush=True)
        return
    if args.spmv_only:
        # copy calibration, then the sparse-product rates at --spmv-workload
        from normalmodes_amd import pevsl as pv
        t_copy = pv.bench_copy(1 << 24, 5)
        print(json.dumps({"copy": {"doubles": 1 << 24, "bytes_per_launch": 16 << 24, "launches": 6,
                                   "us": round(t_copy * 1e6, 2)},
                          "spmv_hbm": spmv_rates(args.spmv_workload, reps=args.steps)}), flush=True)
        return

    from normalmodes_amd import driver, pevsl as pv
    from normalmodes_amd.workloads import build_problem, INTERVALS

    t0 = time.perf_counter()
    P = build_problem(args.workload)
    t_build = time.perf_counter() - t0
    interval = INTERVALS[args.workload]
    log(f"[bench] built {args.workload}: n={P.n} nnzA={P.A.nnz} nnzB={P.B.nnz} in {t_build:.1f}s")
    t0 = time.perf_counter()
    kpart = args.block // args.split if args.block > 1 and args.split > 1 else args.block
    mv = upload_with_layouts(P, args.layouts, sorted({1, args.block, kpart}))
    t_upload = time.perf_counter() - t0
    precision = {"fp64": pv.FP64, "fp32": pv.FP32_FILTER, "fp32a": pv.FP32_FILTER_A,
                 "fp32i": pv.FP32_BSOLVE}[args.precision]

    verbose = max(1, 24 // args.block) if P.n > 500000 else 0   # progress lines of long solves
    for i in range(args.warmup):
        el, r = solve_once(mv, interval, bdeg, precision, 0, args.block, args.split, args.refine, verbose)
        log(f"[bench] warmup {i}: {el:.2f}s nev={r.lam.size} steps={r.lan_steps}")
    times = []
    r = None
    gate_ok = True      # every timed run must meet the gate, not only the last
    for i in range(args.steps):
        el, r = solve_once(mv, interval, bdeg, precision, args.probe_every, args.block, args.split, args.refine,
                           verbose)
        times.append(el)
        gate_ok = gate_ok and bool(r.res2.size > 0 and np.all(r.res2 <= GATE))
        log(f"[bench] step {i}: {el:.2f}s nev={r.lam.size} deg={r.deg} steps={r.lan_steps} "
            f"max res2={np.max(r.res2) if r.res2.size else float('nan'):.2e}")
    T = max(times)
    chk = r.probe["chebstep"]
    achieved = chk["bytes"] / chk["seconds"] / 1e9 if chk["seconds"] > 0 else 0.0
    persistent = pv.chebiter_info(mv.chebB)["persistent"]
    kname = ("cheb_persistent_k" if persistent else
             ("mass3_chebstep" if args.block == 1 else f"block{kpart}_chebstep"))
    if args.refine:
        kname = f"block{kpart}_chebstep32" if kpart > 1 else "chebstep32"
    ktag = f"{kname}:{args.workload}"
    traffic, traffic_src = traffic_from_profiles(ktag)
    line = {
        "metric": METRIC,
        "value": round(T, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "steps_requested": args.requested,
        "ms_per_step": round(T * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64" if precision == pv.FP64 and not args.refine else
                  "f64 with f32 inner Chebyshev passes (f64 residuals)" if precision == pv.FP64 else
                  {pv.FP32_FILTER: "f32 B-solve values, f64 stiffness and recurrence",
                   pv.FP32_FILTER_A: "f32 B-solve and stiffness values in the filter, f64 recurrence",
                   pv.FP32_BSOLVE: "f32 B-solve values and iterates, f64 stiffness and recurrence"}[precision]),
        "data": "synthetic",
        "config": {
            "workload": WORKLOAD_TEXT.get(args.workload, args.workload).format(
                tets=P.mesh.n_elem, n=P.n, a=interval[0], b=interval[1]),
            "refine_passes": args.refine,
            "n_dof": P.n, "nnz_A": P.A.nnz, "nnz_B": P.B.nnz, "n_tets": P.mesh.n_elem,
            "interval": list(interval), "nev": int(r.lam.size), "filter_degree": r.deg,
            "bsolve_degree": mv.bdeg, "lanczos_steps": r.lan_steps,
            "lanbounds_steps": r.lan_steps_bounds, "precision": args.precision, "block": args.block,
            "bsolve_streams": args.block // kpart if args.block > 1 else 1,
            "kernel_layouts": mv.tuned,
            "parallelism": "1 GPU", "global_batch": 1, "seq_len": P.n,
        },
        "gate": {"resid_max": float(np.max(r.res2)) if r.res2.size else None,
                 "resid_ref_max": float(np.max(r.res_ref)) if r.res_ref.size else None,
                 "passed": gate_ok, "threshold": GATE},
        "phases_s": {k: round(v, 4) for k, v in r.timings.items()},
        "setup_s": {"build": round(t_build, 2), "upload_and_autotune": round(t_upload, 3)},
        "roofline": {
            "bound": "hbm",
            "kernel": ("persistent Chebyshev B-solve (all steps in one cooperative launch; bytes = steps x "
                       "one fused step's algorithmic bytes)" if persistent else
                       "fused Chebyshev B-solve step (scalar mass x I3 product + update)" if args.block == 1 else
                       f"fused {kpart}-column Chebyshev B-solve step (scalar mass x I3 product + update on "
                       f"{kpart} interleaved columns)" + (f"; {args.block // kpart} such solves run at once "
                                                           f"on as many streams" if kpart != args.block else "")),
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": traffic_src,
            "bytes_per_launch": chk["bytes"], "mean_launch_us": round(chk["seconds"] * 1e6, 3),
            "samples": chk["samples"],
            # with the B-solve split over streams, the other parts' launches run
            # beside each timed one: their sum is the bandwidth the step draws
            "concurrent_parts": args.block // kpart if args.block > 1 else 1,
            "aggregate_achieved": round(achieved * (args.block // kpart if args.block > 1 else 1), 1),
            "filter_step": r.probe["filterstep"],
        },
    }
    if args.fp64_compare and args.refine:
        try:
            el64, r64 = solve_once(mv, interval, bdeg, pv.FP64, 0, args.block, args.split, 0, verbose)
            line["fp64_plain"] = {
                "seconds": round(el64, 3), "nev": int(r64.lam.size), "lanczos_columns": r64.lan_steps,
                "resid_max": float(np.max(r64.res2)) if r64.res2.size else None,
                "speedup_of_value": round(el64 / T, 3),
                "max_rel_eig_diff_vs_value": float(np.max(np.abs(r64.lam - r.lam) / np.abs(r.lam)))
                if r64.lam.size == r.lam.size else None}
            log(f"[bench] plain fp64: {line['fp64_plain']}")
        except Exception as e:
            line["fp64_plain"] = {"error": str(e)}
    if P.n > 500000:
        try:
            line["filter_cost"] = filter_rates(mv, interval, args.block, args.split, 1e-5)
            log(f"[bench] filter cost: {line['filter_cost']}")
        except Exception as e:
            line["filter_cost"] = {"error": str(e)}
    if args.block > 1:
        try:
            line["reorth_gemm"] = reorth_rate(P.n, min(r.lan_steps, 512), args.block)
        except Exception as e:
            line["reorth_gemm"] = {"error": str(e)}
    if args.single_compare and args.block > 1:
        try:
            el1, r1 = solve_once(mv, interval, bdeg, precision, 0, 1)
            line["single_vector"] = {"seconds": round(el1, 4), "nev": int(r1.lam.size), "lanczos_steps": r1.lan_steps,
                                     "resid_max": float(np.max(r1.res2)) if r1.res2.size else None,
                                     "max_rel_eig_diff_vs_block": float(np.max(np.abs(r1.lam - r.lam) / r.lam))
                                     if r1.lam.size == r.lam.size else None}
            log(f"[bench] single vector: {line['single_vector']}")
        except Exception as e:
            line["single_vector"] = {"error": str(e)}
    if args.fp32_compare:
        try:
            el32, r32 = solve_once(mv, interval, bdeg, pv.FP32_FILTER, 0, args.block, args.split)
            m = min(r32.lam.size, r.lam.size)
            line["fp32_filter"] = {
                "seconds": round(el32, 4), "nev": int(r32.lam.size),
                "max_rel_eig_diff_vs_fp64": float(np.max(np.abs(r32.lam[:m] - r.lam[:m]) / np.abs(r.lam[:m])))
                if m and r32.lam.size == r.lam.size else None,
                "resid_max": float(np.max(r32.res2)) if r32.res2.size else None,
                "note": "FP32 B-solve values; residual floor of the FP32 B-solve is ~1e-6 (not gated)"}
            log(f"[bench] fp32 filter: {line['fp32_filter']}")
        except Exception as e:
            line["fp32_filter"] = {"error": str(e)}
    if args.adjoint:
        try:
            from normalmodes_amd.adjoint import backward
            ra = driver.apply_pevsl(mv, interval=interval, vectors=False, block=args.block, keep=True,
                                    split=args.split if args.block > 1 else 0)
            try:
                backward(ra.handle, P.mesh)   # first call: element data upload and kernel load
                t0 = time.perf_counter()
                g = backward(ra.handle, P.mesh)
                ta = time.perf_counter() - t0
            finally:
                pv.pevsl_finish(ra.handle)
            line["adjoint"] = {"seconds": round(ta, 4), "elements": P.mesh.n_elem, "pairs": int(ra.lam.size),
                               "element_pairs_per_s": round(P.mesh.n_elem * ra.lam.size / ta, 0),
                               "note": "d lambda / d(rho, lambda, mu) per element for every pair, host<->device "
                                       "copies of the element data and the three gradients included"}
            log(f"[bench] adjoint: {line['adjoint']}")
        except Exception as e:
            line["adjoint"] = {"error": str(e)}
    if args.spmv_workload and args.spmv_workload != "none":
        try:
            line["spmv_hbm"] = spmv_rates(args.spmv_workload)
            log(f"[bench] spmv rates: {line['spmv_hbm']}")
        except Exception as e:  # report, but do not lose the headline line
            line["spmv_hbm"] = {"error": str(e)}
    if args.fluid_pieces:
        try:
            line["c2_fluid"] = fluid_pieces(args)
            log(f"[bench] c2 fluid pieces: {line['c2_fluid']}")
        except Exception as e:
            line["c2_fluid"] = {"error": str(e)}
    if args.cpu_seconds > 0:
        try:
            # the CPU restatement runs the reference's single-vector ChebLanNr:
            # scale it by a single-vector solve's filter work
            r_ref = r if args.block == 1 else (r1 if "r1" in locals() else r)
            line["cpu_baseline"] = cpu_baseline(P, mv, r_ref, args.cpu_seconds)
            log(f"[bench] cpu baseline: {line['cpu_baseline']}")
        except Exception as e:
            line["cpu_baseline"] = {"error": str(e)}
    if args.cpu_full:
        try:
            line["cpu_full_solve"] = cpu_full_solve(P, interval, mv.bdeg)
            log(f"[bench] cpu full solve: {line['cpu_full_solve']}")
        except Exception as e:
            line["cpu_full_solve"] = {"error": str(e)}
    mv.free()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
