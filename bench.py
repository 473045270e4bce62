#!/usr/bin/env python3
"""Benchmark of the quasi-static Newton path on one MI355X.

Workload (default): a 128^3-hex cube as 12,582,912 Kuhn tets (2,146,689
nodes, 6,440,067 DOFs), interior jitter, lognormal mu_e, nu = 0.49, x=0
clamped, 25% uniaxial stretch of x=1 over 20 load steps.  A "step" is one
load step solved by Newton to ||R|| <= 1e-8 ||R_0|| (FP64 residual) with FP32
assembly, SpMV, ILU(0) and PCG (the mixed path).  Inputs are synthetic and
already resident in HBM when the timed region starts.

Step plan: the first `--warmup` load steps run once as a throw-away pass,
the state returns to the unloaded configuration (u = 0 at load factor 0), and
load steps 1..K of the path are timed (K = --steps, at most the path's load
steps).  Every library call is synchronous on return, so the wall clock
around the K steps brackets exactly the device work of the timed region.

value = DOFs x (load steps converged) / wall seconds of the timed steps.
After the timed region, on the same load steps from the unloaded state:
  * the plain FP64 GPU path running the same algorithm (extrapolating
    predictor, Galerkin start over recycled corrections and polynomial
    fields, the same fields as an additive coarse level, the first Newton
    iteration's preconditioner kept for the later ones of a step,
    target-aware forcing term) -- the precision-only ratio mixed_over_fp64;
  * the plain FP64 GPU path with the reference's solver settings (predictor
    solve, PCG tol 1e-10, 300-iteration cap), with its cap hits counted;
  * the FP64 CPU oracle on a bounded sample on the host cores (assembly,
    residual, ILU(0) factorization, PCG iterations), extrapolated with the
    operation counts of the same-algorithm FP64 run, and the oracle's FP64
    residual of the final mixed displacement.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "atlas-flexmesh_amd"))

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E
# The timed pass brackets one launch in PROF_EVERY of the kernels launched once
# per PCG iteration (SpMV, the two sweeps, the other PCG kernels) with HIP
# events, and every launch of the other classes: each pair of events puts a
# bubble of a few microseconds into the stream (about 4% of a load step when
# every launch is bracketed).
PROF_EVERY = 16


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_bytes(info, kind, v, launches_per_unit=None):
    """Algorithmic bytes per launch of kernel class `kind` (v = value bytes, 4-byte
    indices).  An assembly or residual is one node-gather launch or one launch per
    colour; launches_per_unit says which (default: one per colour)."""
    n, nnz, ne, nn = info["n"], info["nnz"], info["ne"], info["nn"]
    i = 4
    nlow = (nnz - n) // 2  # strictly lower entries (structurally symmetric pattern)
    if kind == "spmv":
        if info.get("spmv") == "node":
            # values once, one column-node index per 3x3 block, 3 row pointers and a
            # block pointer per node, x once, y once
            return nnz * v + (nnz // 9) * i + (n // 3) * 4 * i + n * 2 * v
        # values + column indices once, row pointers, x once, y once (+ fused p.Ap reads x again, cached)
        return nnz * (v + i) + n * (i + 2 * v)
    chain = info.get("sweep") == "chain"
    if kind == "ilu_fwd":
        if chain:
            # L part values, one dependency word per 3x3 block, rhs in, z out
            # (the PCG update runs in its own kernel on this path)
            return nlow * v + (nlow // 9) * i + n * 2 * v
        # L part values + columns, row pointers, rhs in, z out; fused PCG update reads r, Ap, p, x, writes r, x
        return nlow * (v + i) + n * (i + 2 * v) + n * 6 * v
    if kind == "ilu_bwd":
        if chain:
            # U part values (incl. the diagonal blocks), one word per off-diagonal block, z and r in, y out
            return (nnz - nlow) * v + (nlow // 9) * i + n * 3 * v
        # U part (incl. diagonal) values + columns, row pointers, z and r in, y out
        return (nnz - nlow) * (v + i) + n * (i + 3 * v)
    if kind == "ilu_factor":
        return nnz * (2 * v + i) + n * i
    per = launches_per_unit or max(info["ncolors"], 1)
    if kind == "assemble":
        # the compulsory traffic of one assembly divided over its launches
        return int((ne * (16 + 2 * 8) + nn * 24 + n * 8 + nnz * v + n * 8) / per)
    if kind == "residual":
        return int((ne * (16 + 2 * 8) + nn * 24 + n * 8 + n * 8) / per)
    if kind == "zz":
        # samples: connectivity, material, X and u once, 7 values written and read
        # back, the node -> element lists; node averages and fits written and read
        # back; indicators written
        return ne * (16 + 16 + 2 * 7 * 8 + 16 + 4 + 8) + nn * (24 + 24 + 2 * 48)
    if kind == "coarse":
        # the residual in the PCG precision and the scaled node coordinates
        return n * v + nn * 16
    if kind == "fields":
        # connectivity and the two material values per element, X and u once per
        # node, 21 FP64 outputs per element
        return ne * (16 + 2 * 8 + 21 * 8) + nn * 2 * 24
    return 0


def measured_traffic(kind, prec_name, dims):
    """Calibrated HBM bytes per launch from the committed rocprofv3 PMC summary
    of the full-size configuration 3 workload (None for other workloads)."""
    if tuple(dims) != (128, 128, 128):
        return None
    path = os.path.join(ROOT, "profiles", "config3_128cube_pmc_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 4 and f[0] == "TRAFFIC" and f[1] == kind and f[2] == prec_name:
            return float(f[3])
    return None


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def physical_cores():
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True).stdout
        cores = {l for l in out.splitlines() if l and not l.startswith("#")}
        return max(1, len(cores))
    except Exception:
        return os.cpu_count() or 1


def cpu_baseline(cfg, u, counts, S, budget_its, counts_from):
    """Time the FP64 CPU oracle on a bounded sample of the workload -- one
    tangent assembly with the Dirichlet elimination, one residual, one
    natural-order ILU(0) factorization and budget_its PCG iterations at the
    state u -- and extrapolate one load step with the operation counts per
    load step given in `counts`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    cores = physical_cores()
    O.set_threads(cores)
    X, conn, mu, lam = cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"]
    rp, ci = S.pattern()  # bit-exact with the oracle's pattern (tested)
    colors, nc = S.colors()
    isd = np.zeros(u.size, dtype=np.uint8)
    isd[cfg["dir_dofs"]] = 1
    t = time.perf_counter()
    K, f = O.assemble(X, conn, mu, lam, colors, nc, u, rp, ci)
    O.eliminate(rp, ci, K, isd)
    t_asm = time.perf_counter() - t
    t = time.perf_counter()
    O.assemble(X, conn, mu, lam, colors, nc, u, rp, ci, want_K=False)
    t_res = time.perf_counter() - t
    t = time.perf_counter()
    O.ilu0(rp, ci, K)
    t_fac = time.perf_counter() - t
    # k iterations of the oracle's pcg_solve_ilu0 (which factorizes first) on a
    # right-hand side with the constrained rows zeroed; tol 1e-30 is never met
    k = max(1, int(budget_its))
    b = np.random.default_rng(1).standard_normal(u.size)
    b[isd == 1] = 0.0
    t = time.perf_counter()
    O.pcg(rp, ci, K, b, tol=1e-30, max_iter=k)
    t_it = max(time.perf_counter() - t - t_fac, 1e-9) / k
    t_step = (counts["assemblies"] * t_asm + counts["residuals"] * t_res + counts["factorizations"] * t_fac +
              counts["pcg_iters"] * t_it)
    dofs = u.size
    c = {k: round(v, 2) for k, v in counts.items()}
    sample = (f"one FP64 tangent assembly with Dirichlet elimination ({t_asm:.2f} s), one residual ({t_res:.2f} s), "
              f"one natural-order ILU(0) factorization ({t_fac:.2f} s) and {k} ILU(0)-PCG iterations ({t_it:.3f} s "
              f"each) at {dofs} DOFs on {cores} cores ({cpu_model()}); one load step extrapolated with the "
              f"per-step counts of {counts_from} ({c['assemblies']} assemblies, {c['residuals']} residuals, "
              f"{c['factorizations']} factorizations, {c['pcg_iters']} PCG iterations) = {t_step:.1f} s")
    return {"value": dofs / t_step, "unit": "DOFs/s", "cores": cores, "kind": "port", "sample": sample,
            "seconds_per_load_step": t_step, "t_assemble": t_asm, "t_residual": t_res, "t_factor": t_fac,
            "t_pcg_iter": t_it}


def step_plan(n_steps, steps, warmup):
    """Load steps of the warm-up pass and of the timed pass.

    The warm-up pass runs load steps 1..min(warmup, n_steps) and is thrown
    away; the timed pass restarts from the unloaded state and runs load steps
    1..min(steps, n_steps).  Returns (warm_last, timed_last); 0 = none."""
    if n_steps < 1:
        raise ValueError("a load path needs at least one load step")
    return max(0, min(warmup, n_steps)), max(1, min(steps, n_steps))


def run_steps(S, cfg, prec, first, last, inner_tol, profile, pcg_max=1000, lift=1, **kw):
    """Load steps first..last on the device-resident state; totals of the
    Newton statistics and the wall seconds around them."""
    import flexmesh as fm

    c = fm.configs.newton_cfg(cfg, prec=prec, inner_tol=inner_tol, inner_max=pcg_max, lift=lift, **kw)
    tot = dict(newton_iters=0, pcg_iters=0, assemblies=0, residuals=0, factorizations=0, converged=0,
               pcg_cap_hits=0)
    tot["step_seconds"], tot["r0"], tot["rel_res"] = [], [], []
    worst = 0.0
    if profile:
        S.profile(PROF_EVERY)
        S.profile_reset()
    name = {fm.FM_FP64: "fp64", fm.FM_MIXED: "mixed", fm.FM_MIXED_PC: "mixed-pc"}[prec]
    t0 = time.perf_counter()
    for s in range(first, last + 1):
        rc, st = S.newton_step(s, c, check=False)
        for k in ("newton_iters", "pcg_iters", "assemblies", "residuals", "factorizations", "converged",
                  "pcg_cap_hits"):
            tot[k] += st[k]
        tot["step_seconds"].append(st["seconds"])
        tot["r0"].append(st["r0"])
        tot["rel_res"].append(st["rel_res"])
        worst = max(worst, st["rel_res"])
        log(f"  {name} step {s}: newton {st['newton_iters']} pcg {st['pcg_iters']} (capped solves "
            f"{st['pcg_cap_hits']}) rel_res {st['rel_res']:.2e} {st['seconds']:.2f} s")
    wall = time.perf_counter() - t0
    prof = S.profile_get() if profile else None
    if profile:
        S.profile(False)
    return wall, tot, worst, prof


def oracle_residual(cfg, S, u, load_factor):
    """||R(u)||_2 of the FP64 CPU oracle's residual-only assembly
    (load f_ext - f_int(u), constrained rows zeroed) on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.set_threads(physical_cores())
    rp, ci = S.pattern()
    colors, nc = S.colors()
    t = time.perf_counter()
    _, f = O.assemble(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], colors, nc, u, rp, ci, want_K=False)
    R = -f
    if cfg["f_ext"] is not None:
        R += load_factor * cfg["f_ext"]
    R[cfg["dir_dofs"]] = 0.0
    return float(np.linalg.norm(R)), time.perf_counter() - t


def element_fields_timing(S, info, kind="fields", reps=5):
    """A post-processing pass over the converged state (after the timed region,
    not part of `value`): the element fields or the ZZ estimate (three kernels),
    HIP-event time per call and algorithmic bandwidth."""
    S.profile(True)
    S.profile_reset()
    for _ in range(reps):
        if kind == "fields":
            S.element_fields(None, fields=())
        else:
            S.zz_estimate()
    ms, n = S.profile_get()["fields"]
    S.profile(False)
    us = 1e3 * ms / max(n, 1)
    b = kernel_bytes(info, kind, 8)
    return {"avg_us": us, "launches": n, "bytes_per_launch": b, "GBps": b / (us * 1e-6) / 1e9,
            "frac_of_hbm_peak": b / (us * 1e-6) / 1e9 / PEAK_HBM_GBS}


def micro(S, cfg, info, iters):
    """Per-kernel timings on the tangent at u = 0 with the Dirichlet rows eliminated."""
    import flexmesh as fm

    out = {"config": cfg["k"], "dims": cfg["dims"], "info": info}
    b = np.random.default_rng(0).standard_normal(cfg["X"].size)
    b[cfg["dir_dofs"]] = 0.0
    for name, prec, v in (("mixed", fm.FM_MIXED, 4), ("fp64", fm.FM_FP64, 8)):
        S.profile(True)
        S.profile_reset()
        S.assemble(None, prec, want_K=False)
        S.use_tangent(True)
        t = time.perf_counter()
        _, its, rel = S.pcg(b, tol=1e-30, max_it=iters, prec=prec)
        wall = time.perf_counter() - t
        prof = S.profile_get()
        S.profile(False)
        out[name] = {"pcg_iters": its, "ms_per_pcg_iter": 1e3 * wall / max(its, 1), "relres": rel,
                     "kernels": {k: {"avg_us": 1e3 * ms / n, "launches": n,
                                     "GBps": kernel_bytes(info, k, v, n if k == "assemble" else None)
                                     / (ms / n * 1e-3) / 1e9}
                                 for k, (ms, n) in prof.items() if n}}
        log(f"{name}: {out[name]['ms_per_pcg_iter']:.2f} ms/iter " +
            " ".join(f"{k}={d['avg_us']:.0f}us/{d['GBps']:.0f}GB/s" for k, d in out[name]["kernels"].items()))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed load steps 1..K (at most the path's load steps)")
    ap.add_argument("--warmup", type=int, default=1, help="load steps of the throw-away warm-up pass")
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--dims", type=str, default=None, help="override hex counts, e.g. 64,64,64")
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="PCG forcing term of the inexact Newton solves, the smallest one with --inner-tol-max "
                         "(default 1e-3 with the extrapolating predictor and the Galerkin start, 1e-2 with the "
                         "extrapolating predictor alone, 3e-3 with the predictor solve)")
    ap.add_argument("--inner-tol-max", type=float, default=None,
                    help="largest forcing term of the target-aware rule clamp(0.5 target / ||R||, inner_tol, "
                         "inner_tol_max) (default 0.1 with the extrapolating predictor; 0: a fixed forcing term)")
    ap.add_argument("--coarse", type=int, default=None,
                    help="degree of the polynomial displacement fields in the Galerkin start of every inner "
                         "solve (default 4 with the extrapolating predictor; 0: none)")
    ap.add_argument("--additive", type=int, default=None,
                    help="1: those polynomial fields also form the coarse level of a two-level preconditioner "
                         "in every PCG iteration (default 1 with the extrapolating predictor)")
    ap.add_argument("--order", type=int, default=None,
                    help="degree of the load-step extrapolation through the previous converged states (1..4; "
                         "default 3, cubic through four)")
    ap.add_argument("--slope", type=int, default=None,
                    help="1: the first step's predictor solve (to 1e-5) gives the path's slope at the unloaded "
                         "state, Hermite data of the second and third steps' extrapolation (default 0)")
    ap.add_argument("--factor-steps", type=int, default=None,
                    help="with the lagged preconditioner: load steps one ILU(0) factor serves (default 6 with the "
                         "coarse level; 1: a new factor every step)")
    ap.add_argument("--ritz", type=int, default=None,
                    help="Ritz vectors of each solve (<= 8) in the next step's Galerkin start (default 4 with the "
                         "coarse level)")
    ap.add_argument("--keep-coarse", type=int, default=None,
                    help="1: the steps that keep an earlier factor also keep its coarse operator Z^T K Z (default 0)")
    ap.add_argument("--lag", type=int, default=None,
                    help="1: the later Newton iterations of a load step keep the first iteration's ILU(0) factor "
                         "and coarse operator (a lagged preconditioner; the tangent is re-assembled) (default 1 "
                         "with the coarse level)")
    ap.add_argument("--recycle", type=int, default=None,
                    help="the latest 2x this many Newton corrections of the load path span, with the "
                         "polynomial fields, the Galerkin start of a step's first solve (default 6 with the "
                         "extrapolating predictor; 0: x0 = 0)")
    ap.add_argument("--pcg-max", type=int, default=1000, help="PCG iteration cap of the inexact Newton solves")
    ap.add_argument("--no-fp64", action="store_true", help="skip both FP64 GPU rows")
    ap.add_argument("--fp64-ref-steps", type=int, default=None,
                    help="load steps of the FP64 row with the reference's solver settings (default: all timed)")
    ap.add_argument("--predictor", choices=["extrapolate", "solve"], default="extrapolate",
                    help="load-step predictor: extrapolation of the previous converged steps (quadratic "
                         "through three), or the linear solve with the prescribed increment")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=5.0, help="PCG iterations in the CPU sample")
    ap.add_argument("--micro", type=int, default=0,
                    help="kernel timing only: one assembly, one factorization and this many PCG iterations")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding of this path "
              "is a design note, not a measured configuration", file=sys.stderr)
        sys.exit(2)
    import flexmesh as fm

    dims = tuple(int(v) for v in args.dims.split(",")) if args.dims else None
    t = time.perf_counter()
    cfg = fm.configs.make(args.config, dims=dims)
    log(f"config {args.config} dims {cfg['dims']}: {cfg['X'].size} DOFs, {cfg['conn'].size // 4} tets "
        f"({time.perf_counter() - t:.1f} s)")
    warm_last, K = step_plan(cfg["n_steps"], args.steps, args.warmup)
    if K < args.steps:
        log(f"--steps {args.steps} exceeds the {cfg['n_steps']} load steps of the path: timing load steps 1..{K}")
    S = fm.Solver(0)
    try:
        out = bench(S, cfg, args, warm_last, K)
    finally:
        S.close()
    if out is not None:
        print(json.dumps(out), flush=True)


def bench(S, cfg, args, warm_last, K):
    import flexmesh as fm

    t = time.perf_counter()
    fm.configs.load(S, cfg)
    info = S.info()
    sw = S.sweep_info()
    info["sweep"], info["spmv"] = sw["sweep"], sw["spmv"]
    log(f"set-up (pattern, colouring, schedule, upload) {time.perf_counter() - t:.1f} s: {info} sweeps {sw}")
    mixed = cfg["prec"] == fm.FM_MIXED
    prec = fm.FM_MIXED if mixed else fm.FM_FP64
    if args.micro:
        micro(S, cfg, info, args.micro)
        return None
    # the extrapolating predictor uses the previous converged steps of the
    # path (the unloaded state at load factor 0 among them)
    lift = 2 if args.predictor == "extrapolate" else 1
    # the inner solves: with the extrapolated start (cubic through the last
    # four converged states), every solve starts from the Galerkin projection
    # onto the polynomial displacement fields of degree <= 4 (the smooth
    # modes ILU(0) resolves slowly), the first solve of a step also onto the
    # latest twelve Newton corrections of the path, and the same fields stay
    # in every PCG iteration as the coarse level of a two-level
    # preconditioner z = M^{-1} r + Z (Z^T K Z)^{-1} Z^T r; a solve stops at
    # the target-aware forcing term clamp(0.5 target/||R||, 1e-3, 0.1), so
    # that from the cubic start one Newton iteration reaches the target; a
    # later iteration of a step keeps the first one's factor and coarse
    # operator, and the four Ritz vectors of the slowest modes that the last
    # solve's Lanczos process found join the next step's start; without them
    # the extrapolated start takes 1e-2 throughout, the predictor solve 3e-3
    recycle = args.recycle if args.recycle is not None else (6 if lift == 2 else 0)
    coarse = args.coarse if args.coarse is not None else (4 if lift == 2 else 0)
    additive = args.additive if args.additive is not None else (1 if lift == 2 and coarse else 0)
    inner_tol_max = args.inner_tol_max if args.inner_tol_max is not None else (0.1 if lift == 2 else 0.0)
    if args.inner_tol is not None:
        inner_tol = args.inner_tol
    elif lift == 2:
        inner_tol = 1e-3 if (recycle or inner_tol_max or coarse) else 1e-2
    else:
        inner_tol = 3e-3
    lag = args.lag if args.lag is not None else (1 if additive else 0)
    order = args.order if args.order is not None else 3
    slope = args.slope if args.slope is not None else 0
    fsteps = args.factor_steps if args.factor_steps is not None else (6 if lag else 1)
    kcoarse = args.keep_coarse if args.keep_coarse is not None else 0
    ritz = args.ritz if args.ritz is not None else (4 if additive else 0)
    solver_kw = dict(recycle=recycle, inner_tol_max=inner_tol_max, coarse_degree=coarse, coarse_additive=additive,
                     lag_precond=lag, extrap_order=order, slope_start=slope, factor_steps=fsteps,
                     keep_coarse=kcoarse, ritz=ritz)
    if warm_last:
        log(f"warm-up pass: load steps 1..{warm_last}")
        run_steps(S, cfg, prec, 1, warm_last, inner_tol, False, args.pcg_max, lift, **solver_kw)
    S.set_u(None)  # back to the unloaded state: the timed pass is the load path from its start
    log(f"timed pass: load steps 1..{K} of {cfg['n_steps']}")
    wall, tot, worst, prof = run_steps(S, cfg, prec, 1, K, inner_tol, True, args.pcg_max, lift, **solver_kw)
    u_end = S.get_u()
    dofs = cfg["X"].size
    value = dofs * tot["converged"] / wall
    nx, ny, nz = cfg["dims"]
    out = {
        "metric": "Newton DOFs/s to 1e-8 rel. residual, FP32-mixed vs FP64 GPU vs CPU; HBM GB/s",
        "value": value, "unit": "DOFs/s", "n_gpus": 1, "steps": K, "warmup": warm_last,
        "ms_per_step": 1e3 * wall / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32-mixed" if mixed else "f64",
        "data": "synthetic (seeded mesh jitter and lognormal fields)",
        "config": {"workload": f"config {args.config}: {nx}x{ny}x{nz} hexes x6 Kuhn tets, {dofs} DOFs, "
                               f"load steps 1..{K} of {cfg['n_steps']}", "dofs": dofs, "tets": info["ne"],
                   "nnz": info["nnz"], "predictor": "extrapolate" if lift == 2 else "solve",
                   "inner_tol": inner_tol, "inner_tol_max": inner_tol_max, "recycle": recycle,
                   "coarse_degree": coarse, "coarse_additive": additive, "lag_precond": lag, "extrap_order": order, "slope_start": slope, "factor_steps": fsteps, "keep_coarse": kcoarse, "ritz": ritz,
                   "newton_rtol": cfg["newton_rtol"]},
        "newton": {"converged_steps": tot["converged"], "newton_iters": tot["newton_iters"],
                   "pcg_iters": tot["pcg_iters"], "newton_it_per_s": tot["newton_iters"] / wall,
                   "pcg_it_per_s": tot["pcg_iters"] / wall, "worst_rel_res": worst,
                   "step_seconds": [round(x, 4) for x in tot["step_seconds"]]},
    }
    # roofline of the dominant kernel class
    v = 4 if mixed else 8
    kern = {}
    units = {"assemble": tot["assemblies"], "residual": tot["residuals"]}
    for k, (ms, nl) in prof.items():
        if nl:
            b = kernel_bytes(info, k, v, nl / units[k] if units.get(k) else None)
            kern[k] = {"ms_total": ms, "launches": nl, "avg_us": 1e3 * ms / nl, "bytes_per_launch": b,
                       "GBps": b / (ms / nl * 1e-3) / 1e9 if b else None}
    dom = max(kern, key=lambda k: kern[k]["ms_total"])
    d = kern[dom]
    out["roofline"] = {"kernel": dom, "bound": "hbm", "achieved": d["GBps"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                       "frac": d["GBps"] / PEAK_HBM_GBS,
                       "traffic": measured_traffic(dom, "fp32" if mixed else "fp64", cfg["dims"]),
                       "algorithmic_bytes": d["bytes_per_launch"], "avg_us": d["avg_us"]}
    out["kernels"] = kern
    out["kernel_timing"] = (f"HIP events on the library's stream around one launch in {PROF_EVERY} of spmv, ilu_fwd, "
                            "ilu_bwd, pcg_other and coarse and around every other launch, over the timed pass; "
                            "ms_total = "
                            "mean of the timed launches x launches")
    if "spmv" in kern:
        out["spmv_frac_of_peak"] = kern["spmv"]["GBps"] / PEAK_HBM_GBS
    out["element_fields"] = element_fields_timing(S, info)
    out["zz_estimate"] = element_fields_timing(S, info, "zz")
    counts = {k: tot[k] / K for k in ("assemblies", "residuals", "factorizations", "pcg_iters")}
    counts_from = "the mixed run"
    u64 = None
    if mixed and not args.no_fp64:
        # precision alone: the FP64 path running the same algorithm on the same steps
        S.set_u(None)
        log(f"FP64 GPU, same algorithm: load steps 1..{K}")
        wall64, tot64, worst64, _ = run_steps(S, cfg, fm.FM_FP64, 1, K, inner_tol, False, args.pcg_max, lift,
                                              inexact=1, **solver_kw)
        u64 = S.get_u()
        out["fp64_gpu"] = {"value": dofs * tot64["converged"] / wall64, "unit": "DOFs/s",
                           "ms_per_step": 1e3 * wall64 / K, "newton_iters": tot64["newton_iters"],
                           "pcg_iters": tot64["pcg_iters"], "converged_steps": tot64["converged"],
                           "worst_rel_res": worst64,
                           "algorithm": "the mixed run's: extrapolating predictor, "
                                        + (f"Galerkin start over the corrections of {recycle} previous steps"
                                           + (f" and the polynomial fields of degree <= {coarse}" if coarse else "")
                                           + ", " if (recycle or coarse) else "")
                                        + ("the same fields as the additive coarse level of the preconditioner, "
                                           if coarse and additive else "")
                                        + ("the first iteration's factor and coarse operator kept for the later "
                                           "iterations of a step, " if lag else "")
                                        + (f"one ILU(0) factor for up to {fsteps} load steps, " if lag and fsteps > 1
                                           else "")
                                        + (f"the {ritz} Ritz vectors of the last solve in the start, " if ritz
                                           else "")
                                        + (f"forcing term clamp(0.5 target/||R||, {inner_tol:g}, {inner_tol_max:g})"
                                           if inner_tol_max > inner_tol else f"forcing term {inner_tol:g}")
                                        + ", with the FP64 tangent, SpMV, ILU(0) and PCG"}
        out["mixed_over_fp64"] = value / out["fp64_gpu"]["value"]
        out["mixed_vs_fp64_u_rel_err"] = float(np.abs(u_end - u64).max() / np.abs(u64).max())
        counts = {k: tot64[k] / K for k in ("assemblies", "residuals", "factorizations", "pcg_iters")}
        counts_from = "the same-algorithm FP64 GPU run"
        # the reference's solver settings: predictor solve, PCG tol 1e-10, 300 iterations
        R = K if args.fp64_ref_steps is None else max(1, min(args.fp64_ref_steps, K))
        S.set_u(None)
        log(f"FP64 GPU, reference solver settings: load steps 1..{R}")
        c_ref = fm.configs.newton_cfg(cfg, prec=fm.FM_FP64)
        wallr, totr, worstr, _ = run_steps(S, cfg, fm.FM_FP64, 1, R, c_ref.inner_tol, False, c_ref.inner_max, 1)
        out["fp64_gpu_reference_settings"] = {
            "steps": f"1..{R}", "value": dofs * totr["converged"] / wallr, "unit": "DOFs/s",
            "ms_per_step": 1e3 * wallr / R, "newton_iters": totr["newton_iters"], "pcg_iters": totr["pcg_iters"],
            "linear_solves": totr["factorizations"], "pcg_cap_hits": totr["pcg_cap_hits"],
            "converged_steps": totr["converged"],
            "worst_rel_res": worstr, "pcg_tol": c_ref.pcg_tol, "pcg_cap": c_ref.pcg_max,
            "mixed_over_this": value / (dofs * totr["converged"] / wallr) if totr["converged"] else None}
    if not args.no_cpu:
        cb = cpu_baseline(cfg, u_end, counts, S, args.cpu_budget, counts_from)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["gpu_over_cpu"] = value / cb["value"]
        # the oracle's FP64 residual of the final mixed displacement (and of the FP64 one)
        lf = K / cfg["n_steps"]
        r_mixed, t_res = oracle_residual(cfg, S, u_end, lf)
        chk = {"load_factor": lf, "r0_of_last_step_gpu": tot["r0"][-1], "oracle_norm_R_mixed_u": r_mixed,
               "oracle_rel_res_mixed_u": r_mixed / tot["r0"][-1], "gpu_rel_res_last_step": tot["rel_res"][-1],
               "seconds": t_res}
        if u64 is not None:
            r64, _ = oracle_residual(cfg, S, u64, lf)
            chk["oracle_rel_res_fp64_u"] = r64 / tot["r0"][-1]
        out["oracle_residual_check"] = chk
        log(f"oracle FP64 residual of the final mixed u: {r_mixed:.3e} = {r_mixed / tot['r0'][-1]:.2e} R0")
    return out


if __name__ == "__main__":
    main()
