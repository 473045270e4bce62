"""torch.autograd.Function over the Newton solve and the hand-written adjoint.

    u = newton_solve(mu_e, lam_e, load_scale, problem)

runs the load steps of `problem` on the GPU (through the C ABI) with the
given per-element Lame parameters and load scale, and returns the converged
displacement at load factor 1 as a float64 CPU tensor.  backward() makes one
FP64 PCG adjoint solve with the converged tangent and one per-element
backward kernel (fm_adjoint); it needs only the final state, because the
hyperelastic equilibrium does not depend on the load path.  Load-step states
are checkpointed to a file (not all stored): a problem built with
restart=True re-solves only the last load step from the checkpoint of the
step before it, which is how the finite-difference checks stay cheap.
"""
import os
import tempfile

import numpy as np
import torch

from . import configs
from ._lib import FM_FP64, Solver


class Problem:
    """A configuration bound to one Solver (device context)."""

    def __init__(self, cfg, solver=None, newton=None, restart=False):
        self.cfg = cfg
        self.S = solver or Solver(0)
        configs.load(self.S, cfg)
        self.newton = newton or configs.newton_cfg(cfg, prec=FM_FP64, newton_rtol=1e-12, pcg_tol=1e-13,
                                                   pcg_max=20000, max_newton=40)
        self.restart = restart
        self.ckpt = os.path.join(tempfile.mkdtemp(prefix="fmck"), "prev_step.ckpt")
        self.have_ckpt = False

    def solve(self, mu, lam, scale):
        cfg, S = self.cfg, self.S
        n = cfg["n_steps"]
        restart = self.restart and self.have_ckpt
        if restart:
            # the checkpoint holds the step-(n-1) state of the fields it was
            # written with (a load checks them); the new fields are set on it
            S.set_material(*self.ckpt_fields)
            S.checkpoint_load(self.ckpt)
        S.set_material(mu, lam)
        S.set_bcs(cfg["dir_dofs"], scale * cfg["dir_vals"])
        S.set_load(None if cfg["f_ext"] is None else scale * cfg["f_ext"])
        if restart:
            S.newton_step(n, self.newton)
        else:
            S.set_u(np.zeros(cfg["X"].size))
            for s in range(1, n + 1):
                if s == n:
                    S.checkpoint_save(self.ckpt, s - 1)
                    self.ckpt_fields = (np.array(mu, dtype=np.float64), np.array(lam, dtype=np.float64))
                    self.have_ckpt = True
                S.newton_step(s, self.newton)
        return S.get_u()


class NewtonSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mu, lam, scale, problem):
        s = float(scale)
        u = problem.solve(mu.detach().cpu().numpy().astype(np.float64),
                          lam.detach().cpu().numpy().astype(np.float64), s)
        ctx.problem = problem
        ctx.scale = s
        return torch.from_numpy(u)

    @staticmethod
    def backward(ctx, grad_u):
        g = grad_u.detach().cpu().numpy().astype(np.float64)
        dmu, dlam, dload, _ = ctx.problem.S.adjoint(g, tol=1e-13, max_it=20000)
        # fm_adjoint differentiates w.r.t. a factor t on the (already scaled)
        # loads at t = 1, so d/d(scale) = dload / scale
        return (torch.from_numpy(dmu), torch.from_numpy(dlam),
                torch.tensor(dload / ctx.scale, dtype=torch.float64), None)


def newton_solve(mu, lam, scale, problem):
    return NewtonSolve.apply(mu, lam, scale, problem)
