#!/usr/bin/env python3
"""Benchmark of the quasi-static Newton path on one MI355X.

Workload (default): configuration 3 -- a 128^3-hex cube as 12,582,912 Kuhn
tets (2,146,689 nodes, 6,440,067 DOFs), interior jitter, lognormal mu_e,
nu = 0.49, x=0 clamped, 25% uniaxial stretch of x=1 over 20 load steps.  A
"step" is one load step solved by Newton to ||R|| <= 1e-8 ||R_0|| (FP64
residual) with FP32 assembly, SpMV, ILU(0) and PCG (the mixed path).  Inputs
are synthetic and already resident in HBM when the timed region starts.

value = DOFs x (load steps converged) / wall seconds of the timed steps.
The same steps are then re-run from the same state on the plain FP64 GPU path,
and the FP64 CPU oracle is timed on a bounded sample on the host cores
(assembly, residual, ILU(0) factorization, PCG iterations) and extrapolated
with the operation counts of the FP64 GPU run.

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


def cpu_baseline(cfg, u_start, step, counts, S, budget_s):
    """Time the FP64 CPU oracle on a bounded sample of the same load step and
    extrapolate one load step with the FP64 GPU run's operation counts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    cores = physical_cores()
    O.set_threads(cores)
    X, conn, mu, lam = cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"]
    rp, ci = S.pattern()  # bit-exact with the oracle's pattern (tested)
    colors, nc = S.colors()
    lf = step / cfg["n_steps"]
    u = u_start.copy()
    isd = np.zeros(u.size, dtype=np.uint8)
    isd[cfg["dir_dofs"]] = 1
    t = time.perf_counter()
    K, f = O.assemble(X, conn, mu, lam, colors, nc, u, rp, ci)
    O.eliminate(rp, ci, K, isd)
    t_asm = time.perf_counter() - t
    t = time.perf_counter()
    _, f = O.assemble(X, conn, mu, lam, colors, nc, u, rp, ci, want_K=False)
    t_res = time.perf_counter() - t
    R = -f
    R[isd == 1] = 0.0
    t = time.perf_counter()
    LU, diag = O.ilu0(rp, ci, K)
    t_fac = time.perf_counter() - t
    # k iterations of the oracle's pcg_solve_ilu0 (which factorizes first)
    k = max(1, int(budget_s))
    t = time.perf_counter()
    O.pcg(rp, ci, K, R, tol=1e-30, max_iter=k)
    t_it = max(time.perf_counter() - t - t_fac, 1e-9) / k
    n_it = k
    t_step = (counts["assemblies"] * t_asm + counts["residuals"] * t_res + counts["factorizations"] * t_fac +
              counts["pcg_iters"] * t_it)
    dofs = u.size
    sample = (f"one FP64 assembly ({t_asm:.2f} s), one residual ({t_res:.2f} s), one natural-order ILU(0) "
              f"factorization ({t_fac:.2f} s) and {n_it} PCG iterations ({t_it:.3f} s each) of load step {step} at "
              f"{dofs} DOFs on {cores} cores ({cpu_model()}); one load step extrapolated with the FP64 GPU run's "
              f"counts ({counts['assemblies']} assemblies, {counts['residuals']} residuals, "
              f"{counts['factorizations']} factorizations, {counts['pcg_iters']} PCG iterations) = {t_step:.1f} s")
    return {"value": dofs / t_step, "unit": "DOFs/s", "cores": cores, "kind": "port", "sample": sample,
            "seconds_per_load_step": t_step, "t_assemble": t_asm, "t_residual": t_res, "t_factor": t_fac,
            "t_pcg_iter": t_it}


def run_steps(S, cfg, prec, first, count, inner_tol, profile, pcg_max=1000, lift=1):
    import flexmesh as fm

    c = fm.configs.newton_cfg(cfg, prec=prec, inner_tol=inner_tol, inner_max=pcg_max, lift=lift)
    tot = dict(newton_iters=0, pcg_iters=0, assemblies=0, residuals=0, factorizations=0, converged=0)
    tot["step_seconds"] = []
    worst = 0.0
    if profile:
        S.profile(True)
        S.profile_reset()
    t0 = time.perf_counter()
    for s in range(first, first + count):
        rc, st = S.newton_step(s, c, check=False)
        for k in tot:
            if k != "step_seconds":
                tot[k] += st[k]
        tot["step_seconds"].append(st["seconds"])
        worst = max(worst, st["rel_res"])
        log(f"  {'fp64' if prec == fm.FM_FP64 else 'mixed'} step {s}: newton {st['newton_iters']} pcg "
            f"{st['pcg_iters']} rel_res {st['rel_res']:.2e} {st['seconds']:.2f} s")
    wall = time.perf_counter() - t0
    prof = S.profile_get() if profile else None
    if profile:
        S.profile(False)
    return wall, tot, worst, prof


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
    ap.add_argument("--steps", type=int, default=19, help="timed load steps (after the warm-up steps)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--dims", type=str, default=None, help="override hex counts, e.g. 64,64,64")
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="FP32 PCG tolerance of the mixed path (default 1e-2 with the extrapolating "
                         "predictor, 3e-3 with the predictor solve)")
    ap.add_argument("--pcg-max", type=int, default=1000, help="FP32 PCG iteration cap of the mixed path")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--predictor", choices=["extrapolate", "solve"], default="extrapolate",
                    help="load-step predictor of the mixed path: extrapolation of the previous converged "
                         "steps (quadratic through three), or the linear solve with the prescribed increment")
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
    if args.warmup + args.steps > cfg["n_steps"]:
        raise SystemExit("warmup + steps exceeds the load steps of the configuration")
    S = fm.Solver(0)
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
        return
    # the extrapolating predictor uses the previous converged steps of the
    # path (the zero state at load 0 and the warm-up step count among them)
    lift = 2 if mixed and args.predictor == "extrapolate" else 1
    # the forcing term that minimises the PCG work: a step from the extrapolated
    # start needs two Newton iterations up to 1e-2 (3e-3 from a predictor solve)
    tol_solve = args.inner_tol if args.inner_tol is not None else 3e-3
    inner_tol = args.inner_tol if args.inner_tol is not None else (1e-2 if lift == 2 else 3e-3)
    if args.warmup:
        run_steps(S, cfg, prec, 1, args.warmup, inner_tol, False, args.pcg_max, lift)
    u0 = S.get_u() if args.warmup else np.zeros(cfg["X"].size)
    first = args.warmup + 1
    wall, tot, worst, prof = run_steps(S, cfg, prec, first, args.steps, inner_tol, True, args.pcg_max, lift)
    u_end = S.get_u()
    dofs = cfg["X"].size
    value = dofs * tot["converged"] / wall
    out = {
        "metric": "Newton DOFs/s to 1e-8 rel. residual (FP32-mixed; FP64 GPU and CPU beside it)",
        "value": value, "unit": "DOFs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32-mixed" if mixed else "f64", "data": "synthetic (seeded mesh jitter and lognormal fields)",
        "config": {"workload": f"config {args.config}: {cfg['dims'][0]}x{cfg['dims'][1]}x{cfg['dims'][2]} hexes x6 "
                               f"Kuhn tets, {dofs} DOFs, load steps {first}..{first + args.steps - 1} of "
                               f"{cfg['n_steps']}", "dofs": dofs, "tets": info["ne"], "nnz": info["nnz"],
                   "predictor": "extrapolate" if lift == 2 else "solve", "inner_tol": inner_tol},
        "newton": {"converged_steps": tot["converged"], "newton_iters": tot["newton_iters"],
                   "pcg_iters": tot["pcg_iters"], "newton_it_per_s": tot["newton_iters"] / wall,
                   "worst_rel_res": worst},
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
                       "algorithmic_bytes": d["bytes_per_launch"]}
    out["kernels"] = kern
    if "spmv" in kern:
        out["spmv_frac_of_peak"] = kern["spmv"]["GBps"] / PEAK_HBM_GBS
    out["element_fields"] = element_fields_timing(S, info)
    out["zz_estimate"] = element_fields_timing(S, info, "zz")
    # FP64 GPU path on the same steps from the same state
    if mixed and not args.no_fp64:
        S.set_u(u0)
        wall64, tot64, worst64, _ = run_steps(S, cfg, fm.FM_FP64, first, args.steps, inner_tol, False)
        out["fp64_gpu"] = {"value": dofs * tot64["converged"] / wall64, "unit": "DOFs/s",
                           "ms_per_step": 1e3 * wall64 / args.steps, "newton_iters": tot64["newton_iters"],
                           "pcg_iters": tot64["pcg_iters"], "worst_rel_res": worst64}
        out["mixed_over_fp64"] = value / out["fp64_gpu"]["value"]
        u64 = S.get_u()
        out["mixed_vs_fp64_u_rel_err"] = float(np.abs(u_end - u64).max() / np.abs(u64).max())
        if lift == 2:
            # the mixed path with the predictor solve of the FP64 path, from the same
            # state, on the first three timed steps (beside the secant run's times there)
            k = min(3, args.steps)
            S.set_u(u0)
            wall1, tot1, worst1, _ = run_steps(S, cfg, prec, first, k, tol_solve, False, args.pcg_max)
            out["mixed_predictor_solve"] = {"steps": f"{first}..{first + k - 1}",
                                            "ms_per_step": 1e3 * wall1 / k,
                                            "extrapolate_ms_per_step": 1e3 * sum(tot["step_seconds"][:k]) / k,
                                            "newton_iters": tot1["newton_iters"], "pcg_iters": tot1["pcg_iters"],
                                            "worst_rel_res": worst1}
            # precision alone: both paths with the predictor solve, same steps
            out["mixed_over_fp64_same_predictor"] = sum(tot64["step_seconds"][:k]) / wall1
        counts = {k: tot64[k] / args.steps for k in ("assemblies", "residuals", "factorizations", "pcg_iters")}
    else:
        counts = {k: tot[k] / args.steps for k in ("assemblies", "residuals", "factorizations", "pcg_iters")}
    if not args.no_cpu:
        cb = cpu_baseline(cfg, u0, first, counts, S, args.cpu_budget)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["gpu_over_cpu"] = value / cb["value"]
    S.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
