"""Seeded synthetic problems: the five benchmark configurations of the
quasi-static Newton path, optionally at a reduced size for tests.

  1  unit cube 8^3 hexes, E=1 nu=0.3, x=0 clamped, u_x(x=1) = 0.2 over 10 steps (FP64 CPU reference case)
  2  cube 64^3, interior jitter U(+-0.2h), mu lognormal (median 1, sigma 0.3), nu=0.3, x=1 compressed 30% / 20 steps
  3  cube 128^3 (6,440,067 DOFs), jitter, lognormal mu, nu=0.49, 25% stretch / 20 steps (mixed precision)
  4  cube 48^3, mu_e and lam_e lognormal (seed A), target from seed B, 15% stretch / 10 steps, loss on z=1
  5  beam 8x1x1 as 256x32x32 hexes, E=1 nu=0.45, x=0 clamped, dead transverse traction on x=8 / 40 steps

Meshes come from the library's own box mesher (fm_mesh_box, Kuhn split,
node id k(nx+1)(ny+1) + j(nx+1) + i) and fields from its counter-based
generators, so a configuration is reproduced bit for bit on any host.
"""
import numpy as np

from . import _lib as L

SEED_JITTER = 0x4A49545445520001
SEED_MU = 0x4D55000000000002
SEED_A = 0x4100000000000004
SEED_B = 0x4200000000000004

DIMS = {1: (8, 8, 8), 2: (64, 64, 64), 3: (128, 128, 128), 4: (48, 48, 48), 5: (256, 32, 32)}


def lame(E, nu):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def grid_index(nx, ny, nz):
    NX, NY = nx + 1, ny + 1
    v = np.arange((nx + 1) * (ny + 1) * (nz + 1), dtype=np.int64)
    return v % NX, (v // NX) % NY, v // (NX * NY)


def face_traction(X, conn, on_face, t):
    """Consistent nodal forces of a uniform dead traction t (force per
    reference area) on the boundary triangles whose three nodes satisfy on_face."""
    nn = X.size // 3
    ne = conn.size // 4
    P = X.reshape(3, nn).T
    c = conn.reshape(4, ne).T
    sel = on_face[c].sum(axis=1) >= 3
    f = np.zeros((nn, 3))
    for skip in range(4):
        tri = np.delete(c[sel], skip, axis=1)
        ok = on_face[tri].all(axis=1)
        tri = tri[ok]
        a = 0.5 * np.linalg.norm(np.cross(P[tri[:, 1]] - P[tri[:, 0]], P[tri[:, 2]] - P[tri[:, 0]]), axis=1)
        for v in range(3):
            np.add.at(f, tri[:, v], a[:, None] * np.asarray(t)[None, :] / 3.0)
    return f.ravel()


def make(k, dims=None, L_box=None):
    """Inputs of configuration k (dims overrides the hex counts for tests)."""
    nx, ny, nz = dims if dims is not None else DIMS[k]
    Lx, Ly, Lz = L_box if L_box is not None else ((8.0, 1.0, 1.0) if k == 5 else (1.0, 1.0, 1.0))
    X, conn = L.mesh_box(nx, ny, nz, Lx, Ly, Lz)
    nn, ne = X.size // 3, conn.size // 4
    gi, gj, gk = grid_index(nx, ny, nz)
    if k in (2, 3):
        interior = ((gi > 0) & (gi < nx) & (gj > 0) & (gj < ny) & (gk > 0) & (gk < nz)).astype(np.uint8)
        X = L.mesh_jitter(X, interior, 0.2 * Lx / nx, SEED_JITTER)
    cfg = dict(k=k, dims=(nx, ny, nz), X=X, conn=conn, f_ext=None, loss_nodes=None)
    x0 = np.where(gi == 0)[0]
    x1 = np.where(gi == nx)[0]
    clamp = np.concatenate([3 * x0, 3 * x0 + 1, 3 * x0 + 2])
    if k == 1:
        mu0, lam0 = lame(1.0, 0.3)
        mu, lam = np.full(ne, mu0), np.full(ne, lam0)
        stretch, n_steps = 0.2, 10
        cfg.update(newton_rtol=1e-10, pcg_tol=1e-10, prec=L.FM_FP64)
    elif k == 2:
        mu = L.field_lognormal(ne, 1.0, 0.3, SEED_MU)
        lam = mu * (2 * 0.3 / (1 - 2 * 0.3))
        stretch, n_steps = -0.3, 20
        cfg.update(newton_rtol=1e-8, pcg_tol=1e-10, prec=L.FM_FP64)
    elif k == 3:
        mu = L.field_lognormal(ne, 1.0, 0.3, SEED_MU)
        lam = mu * (2 * 0.49 / (1 - 2 * 0.49))
        stretch, n_steps = 0.25, 20
        cfg.update(newton_rtol=1e-8, pcg_tol=1e-10, prec=L.FM_MIXED)
    elif k == 4:
        mu = L.field_lognormal(ne, 1.0, 0.3, SEED_A)
        lam = L.field_lognormal(ne, 1.5, 0.3, SEED_A + 1)
        cfg["mu_target"] = L.field_lognormal(ne, 1.0, 0.3, SEED_B)
        cfg["lam_target"] = L.field_lognormal(ne, 1.5, 0.3, SEED_B + 1)
        stretch, n_steps = 0.15, 10
        cfg["loss_nodes"] = np.where(gk == nz)[0]
        cfg.update(newton_rtol=1e-12, pcg_tol=1e-12, prec=L.FM_FP64)
    elif k == 5:
        mu0, lam0 = lame(1.0, 0.45)
        mu, lam = np.full(ne, mu0), np.full(ne, lam0)
        stretch, n_steps = 0.0, 40
        on = np.zeros(nn, dtype=bool)
        on[x1] = True
        # transverse (-z) dead traction; magnitude gives > 60 degrees of tip rotation
        cfg["traction"] = 0.0048
        cfg["f_ext"] = face_traction(X, conn, on, np.array([0.0, 0.0, -cfg["traction"]]))
        # the slender beam's bending modes hold ILU(0)-PCG on a plateau of
        # ~4 n_x iterations before it converges, at any tolerance; its tangent
        # (condition number ~1e6-1e7) is too ill-conditioned for FP32 search
        # directions, so the Krylov iteration runs in FP64 with the FP32
        # ILU(0) factor as preconditioner.  The bending modes are smooth: the
        # polynomial coarse level and Galerkin start of the inexact Newton
        # solves remove the plateau (inner_opts, applied to the forcing-term
        # solves of FM_MIXED_PC and FM_MIXED)
        cfg.update(newton_rtol=1e-8, pcg_tol=1e-10, prec=L.FM_MIXED_PC, ls_mode=1, max_newton=60,
                   pcg_max=8 * nx + 400, inner_max=8 * nx + 400, inner_tol=1e-3,
                   inner_opts=dict(lift=2, inner_tol_max=0.1, recycle=6, coarse_degree=4, coarse_additive=1,
                                   lag_precond=1, factor_steps=6, extrap_order=3))
    else:
        raise ValueError(k)
    if stretch != 0.0:
        dofs = np.concatenate([clamp, 3 * x1])
        vals = np.concatenate([np.zeros(clamp.size), np.full(x1.size, stretch * Lx)])
    else:
        dofs, vals = clamp, np.zeros(clamp.size)
    order = np.argsort(dofs, kind="stable")
    cfg.update(mu=mu, lam=lam, dir_dofs=dofs[order].astype(np.uint32), dir_vals=vals[order], n_steps=n_steps)
    return cfg


def load(solver, cfg):
    """Put configuration cfg into a Solver."""
    solver.set_mesh(cfg["X"], cfg["conn"])
    solver.set_material(cfg["mu"], cfg["lam"])
    solver.set_bcs(cfg["dir_dofs"], cfg["dir_vals"])
    solver.set_load(cfg["f_ext"])
    return solver


def newton_cfg(cfg, **kw):
    """The configuration's Newton settings, then the keyword overrides.  A
    configuration's inner_opts (options of the forcing-term inner solves)
    apply when the resulting precision is FM_MIXED or FM_MIXED_PC."""
    c = L.Solver.newton_cfg(n_steps=cfg["n_steps"], prec=cfg["prec"], newton_rtol=cfg["newton_rtol"],
                            pcg_tol=cfg["pcg_tol"])
    for k in ("ls_mode", "max_newton", "pcg_max", "inner_max", "inner_tol"):
        if k in cfg:
            setattr(c, k, cfg[k])
    prec = kw.get("prec", cfg["prec"])
    if prec in (L.FM_MIXED, L.FM_MIXED_PC):
        for k, v in cfg.get("inner_opts", {}).items():
            setattr(c, k, v)
    for k, v in kw.items():
        setattr(c, k, v)
    return c
