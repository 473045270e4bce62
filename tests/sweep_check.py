"""Run in a child process by the GPU tests: ILU(0) apply of an elastic tangent
through the C ABI against the oracle, with the sweep kind the environment
selects (FM_SWEEP=level forces the level-scheduled sweeps; FM_SWEEP_LOADER=0
or 1 selects the one-wave or the two-wave chain-sweep kernel).  With an output
path it also writes the sweep results of three meshes (ILU(0) applies and PCG
solves in both precisions) for the parent test to compare bit for bit.

  python tests/sweep_check.py level|chain [OUT.npz]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [os.path.join(ROOT, "atlas-flexmesh_amd"), os.path.join(ROOT, "oracle"), HERE]
import flexmesh as fm  # noqa: E402
import oracle as O  # noqa: E402
from test_gpu_parity import elastic_matrix, rel  # noqa: E402


def main():
    want = sys.argv[1]
    S = fm.Solver(0)
    rp, ci, va = elastic_matrix(O, fm, dims=(7, 6, 5))
    S.set_matrix(rp, ci, va)
    kind = S.sweep_info()["sweep"]
    assert kind == want, kind
    LUo, diag = O.ilu0(rp, ci, va)
    x = np.random.default_rng(1).standard_normal(rp.size - 1)
    zo = O.ilu0_apply(rp, ci, LUo, diag, x)
    e64 = rel(S.ilu0_apply(x), zo)
    LU32 = S.ilu0_factor(fm.FM_MIXED)
    n = rp.size - 1
    dpos = np.array([rp[i] + np.searchsorted(ci[rp[i]:rp[i + 1]], i) for i in range(n)])
    z32o = O.ilu0_apply(rp, ci, LU32, LU32[dpos], x.astype(np.float32).astype(np.float64))
    e32 = rel(S.ilu0_apply(x, fm.FM_MIXED), z32o)
    xo, itso, _ = O.pcg(rp, ci, va, x, tol=1e-10, max_iter=500)
    xg, its, _ = S.pcg(x, tol=1e-10, max_it=500)
    print(f"{kind}: ilu fp64 {e64:.2e} fp32 {e32:.2e} pcg its {its} vs {itso} x {rel(xg, xo):.2e}")
    assert e64 < 1e-12 and e32 < 1e-4 and abs(its - itso) <= 1 and rel(xg, xo) < 1e-9
    if len(sys.argv) > 2:
        out = {}
        # a cube, a beam (one bundle, long chains) and a mesh whose bundles
        # have fewer steps than the kernels' unrolled blocks
        for dims in ((7, 6, 5), (20, 3, 3), (2, 2, 1)):
            rp, ci, va = elastic_matrix(O, fm, dims=dims)
            S.set_matrix(rp, ci, va)
            tag = "x".join(map(str, dims))
            out[f"{tag}_kind"] = np.array([S.sweep_info()["sweep"] == "chain"])
            x = np.random.default_rng(2).standard_normal(rp.size - 1)
            out[f"{tag}_z64"] = S.ilu0_apply(x)
            S.ilu0_factor(fm.FM_MIXED)
            out[f"{tag}_z32"] = S.ilu0_apply(x, fm.FM_MIXED)
            for name, prec, tol in (("pcg64", fm.FM_FP64, 1e-10), ("pcg32", fm.FM_MIXED, 1e-5)):
                xs, its, relres = S.pcg(x, tol=tol, max_it=500, prec=prec)
                out[f"{tag}_{name}"] = np.append(xs, [its, relres])
        np.savez(sys.argv[2], **out)
    S.close()
    print(f"SWEEP OK {kind}")


if __name__ == "__main__":
    main()
