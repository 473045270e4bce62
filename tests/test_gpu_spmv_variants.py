"""The two node-blocked SpMV kernels (direct gathers, and LDS-staged value
runs) against each other and against the oracle's CSR product
(FEM/fem_types.hpp:310-319), through the C ABI.

The staged kernel reads the same values in the same order and sums them in the
same order, so the two must agree bit for bit, in FP64 and FP32, in the plain
product and in the PCG's product with the fused p.Ap partials.  The cases cover
a partial last wave, a node with more blocks than its 16-lane group has lanes,
and a value run longer than a wave's LDS slice (read directly).

Tolerances against the oracle as in test_gpu_parity.py: FP64 <= 1e-13 of the
largest |y|, FP32 <= 2e-6.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def run_child(lds, out):
    env = dict(os.environ)
    env["FM_SPMV_LDS"] = "1" if lds else "0"
    r = subprocess.run([sys.executable, os.path.join(HERE, "spmv_check.py"), str(out)], capture_output=True,
                       text=True, timeout=300, env=env)
    assert r.returncode == 0 and "SPMV OK" in r.stdout, r.stdout + r.stderr
    return dict(np.load(out))


def test_staged_and_direct_spmv_same_bits_and_match_oracle(O, tmp_path):
    sys.path.insert(0, HERE)
    import spmv_check

    staged = run_child(True, tmp_path / "staged.npz")
    direct = run_child(False, tmp_path / "direct.npz")
    assert staged.keys() == direct.keys() and len(staged) == 8
    for k in staged:
        assert np.array_equal(staged[k], direct[k]), k
    for name, X, conn, mu, lam, dirs, u, x in spmv_check.problems():
        nn = X.size // 3
        rp, ci = O.pattern(nn, conn)
        colors, nc = O.color(nn, conn)
        K, _ = O.assemble(X, conn, mu, lam, colors, nc, u, rp, ci)
        isd = np.zeros(3 * nn, dtype=np.uint8)
        isd[dirs] = 1
        O.eliminate(rp, ci, K, isd)
        yo = O.spmv(rp, ci, K, x)
        scale = np.abs(yo).max()
        assert np.abs(staged[f"{name}_y64"] - yo).max() <= 1e-13 * scale, name
        assert np.abs(staged[f"{name}_y32"] - yo).max() <= 2e-6 * scale, name
    # six PCG iterations ran in both precisions (no early stop at tol 1e-30)
    assert staged["cube_pcg64"][-2] == 6 and staged["cube_pcg32"][-2] == 6
