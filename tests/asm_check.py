"""Run in a child process by the GPU tests: assembly, residual and one Newton
step through the C ABI with the assembly path the environment selects
(FM_ASM=color forces the colour-launch assembly; it is read once per process).
Writes the results to an .npz for the parent test to compare.

  python tests/asm_check.py OUT.npz
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [os.path.join(ROOT, "atlas-flexmesh_amd"), HERE]
import flexmesh as fm  # noqa: E402


def problem():
    """A jittered 5x4x6 cube with lognormal mu, x=1 stretched (config 3 form)
    and a fixed small displacement field."""
    cfg = fm.configs.make(3, dims=(5, 4, 6))
    cfg["n_steps"] = 4
    u = np.random.default_rng(11).uniform(-0.02, 0.02, cfg["X"].size)
    return cfg, u


def compute(S):
    cfg, u = problem()
    fm.configs.load(S, cfg)
    out = {}
    out["f64"], out["K64"] = S.assemble(u, fm.FM_FP64)
    out["f32"], out["K32"] = S.assemble(u, fm.FM_MIXED)
    out["R"], _ = S.residual(u, 0.5)
    # a load step with the predictor solve: the lift (K times the prescribed
    # increment) and the Dirichlet masking of the tangent assembly
    c = fm.configs.newton_cfg(cfg, prec=fm.FM_FP64, newton_rtol=1e-10, pcg_tol=1e-12, pcg_max=2000)
    rc, st = S.newton_step(1, c)
    out["u1"] = S.get_u()
    out["newton1"] = np.array([st["newton_iters"], st["converged"]])
    cm = fm.configs.newton_cfg(cfg, inner_tol=1e-4)
    S.set_u(None)
    rc, st = S.newton_step(1, cm)
    out["u1_mixed"] = S.get_u()
    return out


def main():
    S = fm.Solver(0)
    try:
        out = compute(S)
    finally:
        S.close()
    np.savez(sys.argv[1], **out)
    print(f"ASM OK {os.environ.get('FM_ASM', 'node')}")


if __name__ == "__main__":
    main()
