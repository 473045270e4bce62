"""Per-operation wall time of the FP64 CPU oracle against the compiled
reference (oracle/_ref/libref.so, built by oracle/build_ref.sh from the
reference sources) on the same inputs and threads, at sizes where both run.

TEST INFRASTRUCTURE ONLY: it shows that the CPU baseline bench.py times (the
oracle) is no slower per operation than the reference's own code, which cannot
run the benchmark sizes (its ILU(0) factorization is O(n^2)).

    python oracle/compare_speed.py [--threads 8] [--sizes 8,12,16]

Prints a markdown table: seconds per operation, reference / oracle.
Assembly and residual: the reference on its own pattern
(build_csr_sparsity_pattern_rigorous), the oracle on the intended pattern of
the same Kuhn box mesh.  SpMV, ILU(0) and PCG: both on the oracle's tangent
with the Dirichlet rows of x = 0 eliminated (the same CSR arrays).
"""
import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import oracle as O  # noqa: E402

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def ref_lib():
    path = os.path.join(HERE, "_ref", "libref.so")
    if not os.path.exists(path):
        raise SystemExit("oracle/_ref/libref.so missing: run oracle/build_ref.sh first")
    L = C.CDLL(path)
    L.ref_time_assemble.argtypes = [C.c_uint32, f64p, C.c_uint32, u32p, C.c_double, C.c_double, f64p, C.c_int,
                                    C.c_int]
    L.ref_time_assemble.restype = C.c_double
    L.ref_time_spmv.argtypes = [C.c_uint32, u32p, u32p, f64p, f64p, f64p, C.c_int]
    L.ref_time_spmv.restype = C.c_double
    L.ref_ilu0.argtypes = [C.c_uint32, u32p, u32p, f64p, f64p, f64p, f64p]
    L.ref_pcg.argtypes = [C.c_uint32, u32p, u32p, f64p, f64p, f64p, C.c_double, C.c_int]
    L.ref_pcg.restype = C.c_int
    return L


def best(f, reps):
    t = 1e300
    for _ in range(reps):
        t0 = time.perf_counter()
        f()
        t = min(t, time.perf_counter() - t0)
    return t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--sizes", type=str, default="8,12,16")
    ap.add_argument("--pcg-its", type=int, default=20)
    args = ap.parse_args()
    L = ref_lib()
    O.set_threads(args.threads)  # one OpenMP runtime: the reference's parallel loops use it too
    E, nu = 1.0, 0.3
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    print(f"| hexes | DOFs | operation | reference (s) | oracle (s) | reference / oracle |")
    print("|---|---|---|---|---|---|")
    for N in (int(v) for v in args.sizes.split(",")):
        X, conn = O.box_mesh(N, N, N)
        nn, ne = X.size // 3, conn.size // 4
        n = 3 * nn
        u = np.random.default_rng(N).uniform(-0.01, 0.01, n)
        rp, ci = O.pattern(nn, conn)
        colors, nc = O.color(nn, conn)
        m, lm = np.full(ne, mu), np.full(ne, lam)
        rows = []
        t_ref = L.ref_time_assemble(nn, X, ne, conn, E, nu, u, 3, 0)
        t_orc = best(lambda: O.assemble(X, conn, m, lm, colors, nc, u, rp, ci), 3)
        rows.append(("tangent + force assembly", t_ref, t_orc))
        t_ref = L.ref_time_assemble(nn, X, ne, conn, E, nu, u, 3, 1)
        t_orc = best(lambda: O.assemble(X, conn, m, lm, colors, nc, u, rp, ci, want_K=False), 3)
        rows.append(("residual-only assembly", t_ref, t_orc))
        K, _ = O.assemble(X, conn, m, lm, colors, nc, u, rp, ci)
        gi = np.arange(nn) % (N + 1)
        isd = np.zeros(n, np.uint8)
        for d in range(3):
            isd[3 * np.where(gi == 0)[0] + d] = 1
        O.eliminate(rp, ci, K, isd)
        x = np.random.default_rng(1).standard_normal(n)
        y = np.zeros(n)
        t_ref = L.ref_time_spmv(n, rp, ci, K, x, y, 5)
        t_orc = best(lambda: O.spmv(rp, ci, K, x), 5)
        rows.append(("SpMV", t_ref, t_orc))
        Lv, Uv, dg = np.zeros(K.size), np.zeros(K.size), np.zeros(n)
        t_fac_ref = best(lambda: L.ref_ilu0(n, rp, ci, K, Lv, Uv, dg), 1)
        t_fac_orc = best(lambda: O.ilu0(rp, ci, K), 3)
        rows.append(("ILU(0) factorization", t_fac_ref, t_fac_orc))
        b = x.copy()
        b[isd == 1] = 0.0
        k = args.pcg_its
        xr = np.zeros(n)
        t_ref = (best(lambda: L.ref_pcg(n, rp, ci, K, b, xr, 1e-30, k), 1) - t_fac_ref) / k
        t_orc = (best(lambda: O.pcg(rp, ci, K, b, tol=1e-30, max_iter=k), 1) - t_fac_orc) / k
        rows.append(("one PCG iteration", t_ref, t_orc))
        for name, a, o in rows:
            print(f"| {N}^3 | {n} | {name} | {a:.3g} | {o:.3g} | {a / o:.2f} |")
        sys.stdout.flush()


if __name__ == "__main__":
    main()
