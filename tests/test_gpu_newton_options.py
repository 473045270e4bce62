"""Newton-loop options of the C ABI (the NonlinearNewtonSolver configuration
members, FEM/nonlinear_solver.hpp:23-64, and the FP64 inexact mode) and the
boundary's state handling: per-iteration histories, the unloaded-state reset,
residual-only assembly without side effects, new coordinates on a kept
topology, and repeated external matrices with one pattern.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


@pytest.fixture(scope="module")
def fm():
    import flexmesh

    flexmesh.lib()
    return flexmesh


@pytest.fixture
def S(fm):
    s = fm.Solver(0)
    yield s
    s.close()


def cube(fm, n_steps=4, dims=(5, 5, 5)):
    cfg = fm.configs.make(3, dims=dims)
    cfg["n_steps"] = n_steps
    return cfg


def oracle(O, cfg, **kw):
    return O.NewtonProblem(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], cfg["dir_dofs"], cfg["dir_vals"],
                           cfg["f_ext"], cfg["n_steps"], **kw)


def test_history_residuals_steps_and_energies(fm, S, O):
    cfg = cube(fm)
    fm.configs.load(S, cfg)
    c = fm.configs.newton_cfg(cfg, prec=fm.FM_FP64, record_energy=1, pcg_tol=1e-12, pcg_max=2000)
    rc, st = S.newton_step(1, c)
    res, alpha, en = S.newton_history()
    # r0, the state after the predictor solve, one entry per Newton iteration
    assert res.size == st["newton_iters"] + 2 and alpha.size == res.size - 1 and en.size == res.size
    assert res[0] == st["r0"] and abs(res[-1] - st["rel_res"] * st["r0"]) <= 1e-15 * st["r0"]
    assert alpha[0] == 1.0 and ((alpha > 0) & (alpha <= 1)).all()
    P = oracle(O, cfg)
    lf = 1 / cfg["n_steps"]
    u = S.get_u()
    assert abs(en[-1] - P.energy(lf, u)) <= 1e-12 * abs(P.energy(lf, u))
    u0 = np.zeros_like(u)
    u0[cfg["dir_dofs"]] = lf * cfg["dir_vals"]
    assert abs(en[0] - P.energy(lf, u0)) <= 1e-12 * abs(P.energy(lf, u0))
    # without record_energy the energies are NaN and nothing else changes
    S.set_u(None)
    rc, st2 = S.newton_step(1, fm.configs.newton_cfg(cfg, prec=fm.FM_FP64, pcg_tol=1e-12, pcg_max=2000))
    res2, alpha2, en2 = S.newton_history()
    assert np.isnan(en2).all() and np.array_equal(res2, res) and np.array_equal(S.get_u(), u)


def test_fp64_inexact_runs_the_mixed_algorithm(fm, S, O):
    cfg = cube(fm)
    fm.configs.load(S, cfg)
    P = oracle(O, cfg, pcg_tol=1e-12, pcg_max=2000)
    uo = np.zeros(cfg["X"].size)
    for s in range(1, 5):
        P.step(s, uo)
    its = {}
    for inexact in (0, 1):
        S.set_u(None)
        c = fm.configs.newton_cfg(cfg, prec=fm.FM_FP64, lift=2, inexact=inexact, inner_tol=1e-2, pcg_tol=1e-12,
                                  pcg_max=2000)
        its[inexact] = 0
        for s in range(1, 5):
            rc, st = S.newton_step(s, c)
            assert st["converged"] and st["rel_res"] <= 1e-8 and st["pcg_cap_hits"] == 0
            its[inexact] += st["pcg_iters"]
        # both sides stop at ||R|| <= 1e-8 ||R_0||: their displacements differ by
        # up to the tangent's condition number times that (nu = 0.49)
        assert rel(S.get_u(), uo) < 1e-6
    # the forcing term 1e-2 takes far fewer PCG iterations than pcg_tol 1e-12
    assert its[1] < 0.6 * its[0], its


def test_reset_to_unloaded_state_repeats_the_load_path(fm, S):
    cfg = cube(fm)
    fm.configs.load(S, cfg)
    c = fm.configs.newton_cfg(cfg, lift=2, inner_tol=1e-2)
    runs = []
    for _ in range(2):
        S.set_u(None)
        assert not S.get_u().any()
        stats = [S.newton_step(s, c)[1] for s in range(1, 5)]
        runs.append((S.get_u(), [(t["newton_iters"], t["pcg_iters"], t["factorizations"]) for t in stats]))
    assert np.array_equal(runs[0][0], runs[1][0]) and runs[0][1] == runs[1][1]
    # step 1 knows the unloaded state only (predictor solve), step 3 on extrapolate
    assert runs[0][1][0][2] == runs[0][1][0][0] + 1 and runs[0][1][2][2] <= runs[0][1][2][0]


def test_absolute_tolerance_energy_test_and_reuse(fm, S):
    cfg = cube(fm)
    fm.configs.load(S, cfg)
    c = fm.configs.newton_cfg(cfg, prec=fm.FM_FP64, newton_rtol=1e-10, pcg_tol=1e-12, pcg_max=2000)
    rc, st = S.newton_step(1, c)
    u1 = S.get_u()
    # the converged state passes tol_absolute at once: no predictor, no iteration
    ca = fm.configs.newton_cfg(cfg, prec=fm.FM_FP64, newton_rtol=1e-10, tol_absolute=1e-6)
    rc, st = S.newton_step(1, ca)
    res, _, _ = S.newton_history()
    assert rc == 0 and st["converged"] and st["newton_iters"] == 0 and st["factorizations"] == 0 and res.size == 1
    assert np.array_equal(S.get_u(), u1)
    # a loose energy test stops after the first iteration
    S.set_u(None)
    ce = fm.configs.newton_cfg(cfg, prec=fm.FM_FP64, newton_rtol=1e-14, energy_tol=1e3, pcg_tol=1e-12,
                               pcg_max=2000)
    rc, st = S.newton_step(1, ce)
    _, _, en = S.newton_history()
    assert st["converged"] and st["newton_iters"] == 1 and st["rel_res"] > 1e-14 and not np.isnan(en).any()
    # reuse_tangent: the predictor solve's tangent, then one more assembly and
    # factorization for all of the step's Newton iterations
    S.set_u(None)
    cr = fm.configs.newton_cfg(cfg, prec=fm.FM_FP64, newton_rtol=1e-10, pcg_tol=1e-12, pcg_max=2000,
                               reuse_tangent=1, max_newton=100)
    rc, st = S.newton_step(1, cr)
    assert st["converged"] and st["assemblies"] == 2 and st["factorizations"] == 2 and st["newton_iters"] > 1
    assert rel(S.get_u(), u1) < 1e-7  # both at rtol 1e-10, amplified by the condition number


@pytest.mark.parametrize("beta,alpha0", [(0.0, 1.0), (1.0, 1.0), (0.5, 0.0)])
def test_bad_line_search_settings_are_rejected(fm, S, beta, alpha0):
    cfg = cube(fm)
    fm.configs.load(S, cfg)
    c = fm.configs.newton_cfg(cfg, ls_beta=beta, ls_alpha0=alpha0)
    rc, st = S.newton_step(1, c, check=False)
    assert rc == fm.FM_EINVAL


def test_line_search_beta_changes_the_backtracking(fm, S):
    # a 60% compression in one step without predictor backtracks; beta sets the steps tried
    cfg = cube(fm, n_steps=1, dims=(4, 4, 4))
    cfg["dir_vals"] = cfg["dir_vals"] * -2.4
    fm.configs.load(S, cfg)
    seen = {}
    for beta in (0.5, 0.25):
        S.set_u(None)
        c = fm.configs.newton_cfg(cfg, prec=fm.FM_FP64, lift=0, ls_beta=beta, pcg_tol=1e-12, pcg_max=2000,
                                  max_newton=60)
        S.newton_step(1, c, check=False)
        _, alpha, _ = S.newton_history()
        seen[beta] = set(np.round(np.log(alpha) / np.log(beta), 9))
    assert all(float(k).is_integer() for k in seen[0.5] | seen[0.25])
    assert any(k > 0 for k in seen[0.5])


def test_residual_only_leaves_the_constraints(fm, S, O):
    cfg = cube(fm, dims=(4, 3, 5))
    fm.configs.load(S, cfg)
    u = np.random.default_rng(5).uniform(-0.02, 0.02, cfg["X"].size)
    nn = cfg["X"].size // 3
    rp, ci = O.pattern(nn, cfg["conn"])
    colors, nc = O.color(nn, cfg["conn"])
    _, fo = O.assemble(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], colors, nc, u, rp, ci, want_K=False)
    f = S.internal_force(u)
    assert rel(f, fo) < 1e-12 and np.abs(f[cfg["dir_dofs"]]).max() > 0
    R, _ = S.residual(u, 0.5)
    assert not R[cfg["dir_dofs"]].any()  # the Dirichlet set is still there
    Ro = -fo
    Ro[cfg["dir_dofs"]] = 0.0
    assert rel(R, Ro) < 1e-12


def test_new_coordinates_on_a_kept_topology(fm, O):
    cfg = cube(fm, dims=(4, 4, 3))
    X2 = fm.mesh_jitter(cfg["X"], np.ones(cfg["X"].size // 3, np.uint8), 0.02, 99)
    u = np.random.default_rng(6).uniform(-0.02, 0.02, cfg["X"].size)
    A, B = fm.Solver(0), fm.Solver(0)
    try:
        fm.configs.load(A, cfg)
        ra = A.pattern()
        fa, Ka = A.assemble(u)
        A.set_coords(X2)
        f2, K2 = A.assemble(u)
        cfg2 = dict(cfg, X=X2)
        fm.configs.load(B, cfg2)
        fb, Kb = B.assemble(u)
        assert np.array_equal(f2, fb) and np.array_equal(K2, Kb) and not np.array_equal(fa, f2)
        assert all(np.array_equal(x, y) for x, y in zip(A.pattern(), ra))
        # a Newton step on the moved nodes matches the fresh context's
        c = fm.configs.newton_cfg(cfg, prec=fm.FM_FP64)
        A.set_u(None)
        A.newton_step(1, c)
        B.newton_step(1, c)
        assert np.array_equal(A.get_u(), B.get_u())
        with pytest.raises(fm.FlexMeshError):
            A.set_coords(X2[:-3])
    finally:
        A.close()
        B.close()


def test_repeated_matrix_pattern_uploads_values(fm, S, O):
    # configuration 2 form (nu = 0.3): PCG iteration counts are stable to +-1
    cfg = fm.configs.make(2, dims=(4, 4, 4))
    nn = cfg["X"].size // 3
    rp, ci = O.pattern(nn, cfg["conn"])
    colors, nc = O.color(nn, cfg["conn"])
    rng = np.random.default_rng(8)
    # caller storage order: each row's columns reversed (exercises the kept sort)
    r64 = rp.astype(np.int64)
    rev = np.concatenate([np.arange(r64[r + 1] - 1, r64[r] - 1, -1) for r in range(rp.size - 1)])
    b = rng.standard_normal(rp.size - 1)
    for k in range(3):
        u = rng.uniform(-0.02, 0.02, 3 * nn)
        K, _ = O.assemble(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], colors, nc, u, rp, ci)
        isd = np.zeros(3 * nn, np.uint8)
        isd[cfg["dir_dofs"]] = 1
        O.eliminate(rp, ci, K, isd)
        S.set_matrix(rp, ci[rev], K[rev])
        x, its, relres = S.pcg(b, tol=1e-10, max_it=2000)
        xo, itso, _ = O.pcg(rp, ci, K, b, tol=1e-10, max_iter=2000)
        assert abs(its - itso) <= 1 and rel(x, xo) < 1e-9
        assert rel(S.spmv(xo), O.spmv(rp, ci, K, xo)) < 1e-13
        # a context that plans this matrix from scratch gives the same bits
        F = fm.Solver(0)
        try:
            F.set_matrix(rp, ci[rev], K[rev])
            xf, itsf, _ = F.pcg(b, tol=1e-10, max_it=2000)
        finally:
            F.close()
        assert itsf == its and np.array_equal(xf, x)


@pytest.mark.parametrize("prec,coarse,additive,lag,order,recycle,slope",
                         [("mixed", 0, 0, 0, 2, 2, 0), ("fp64", 0, 0, 0, 2, 2, 0), ("mixed", 3, 0, 0, 2, 2, 0),
                          ("fp64", 2, 0, 0, 2, 2, 0), ("mixed", 3, 1, 0, 2, 2, 0), ("fp64", 2, 1, 0, 2, 2, 0),
                          ("mixed", 4, 1, 0, 2, 2, 0), ("fp64", 1, 1, 0, 2, 2, 0), ("mixed", 3, 1, 1, 2, 2, 0),
                          ("fp64", 3, 1, 1, 2, 2, 0), ("mixed", 0, 0, 1, 2, 2, 0), ("mixed", 4, 1, 1, 3, 6, 0),
                          ("fp64", 4, 1, 1, 3, 6, 0), ("mixed", 2, 1, 1, 1, 4, 0), ("mixed", 4, 1, 1, 4, 6, 0),
                          ("mixed", 4, 1, 1, 3, 6, 1), ("fp64", 4, 1, 1, 3, 6, 1), ("mixed", 4, 1, 2, 3, 6, 0),
                          ("mixed", 4, 1, 3, 3, 6, 0), ("fp64", 3, 1, 4, 3, 6, 0)])
def test_galerkin_start_and_target_forcing(fm, S, O, prec, coarse, additive, lag, order, recycle, slope):
    """recycle = 2 with the target-aware forcing term (inner_tol 1e-3, at most
    0.1), and with the polynomial fields of degree <= coarse in the Galerkin
    start (and, with additive, in every PCG iteration as the coarse level of
    a two-level preconditioner; with lag, the later Newton iterations of a
    step keep the first one's ILU(0) factor and coarse operator, and with
    lag > 1 one factor serves up to lag load steps, with lag > 2 together
    with its coarse operator; order: the
    degree of the load-step extrapolation, recycle: half the corrections in
    the first solve's start, slope: the predictor solve's slope as Hermite
    data of the second and third steps' extrapolation): every step
    still reaches ||R|| <= 1e-8 ||R_0||, the final state
    matches the FP64 oracle's, and from the third step on the steps take
    fewer PCG iterations than with the same forcing and x0 = 0.
    A reset to the unloaded state clears the start basis with the rest of the
    load-path history: the path then repeats bit for bit."""
    cfg = cube(fm, n_steps=8, dims=(8, 8, 8))
    fm.configs.load(S, cfg)
    P = oracle(O, cfg, pcg_tol=1e-12, pcg_max=4000)
    uo = np.zeros(cfg["X"].size)
    for s in range(1, 9):
        P.step(s, uo)
    kw = dict(prec=fm.FM_MIXED) if prec == "mixed" else dict(prec=fm.FM_FP64, inexact=1)
    plain = fm.configs.newton_cfg(cfg, lift=2, inner_tol=1e-3, inner_tol_max=0.1, recycle=0, **kw)
    rec = fm.configs.newton_cfg(cfg, lift=2, inner_tol=1e-3, inner_tol_max=0.1, recycle=recycle,
                                coarse_degree=coarse, coarse_additive=additive, lag_precond=min(lag, 1),
                                factor_steps=lag, keep_coarse=int(lag >= 3), extrap_order=order,
                                slope_start=slope, **kw)
    runs = {}
    for name, c in (("plain", plain), ("rec", rec), ("rec again", rec)):
        S.set_u(None)
        stats = []
        for s in range(1, 9):
            rc, st = S.newton_step(s, c)
            assert st["converged"] and st["rel_res"] <= 1e-8 and st["pcg_cap_hits"] == 0, (name, s, st)
            stats.append((st["newton_iters"], st["pcg_iters"], st["rel_res"]))
        runs[name] = (S.get_u(), stats)
        # both stop at ||R|| <= 1e-8 ||R_0|| (nu = 0.49): up to the tangent's
        # condition number times that apart
        assert rel(runs[name][0], uo) < 1e-6, name
    assert np.array_equal(runs["rec"][0], runs["rec again"][0]) and runs["rec"][1] == runs["rec again"][1]
    later = lambda name: sum(t[1] for t in runs[name][1][2:])
    assert later("rec") < later("plain"), (runs["rec"][1], runs["plain"][1])


# monomials of the polynomial fields (exponents of x, y, z), in the kernels' order
MONOS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (0, 1, 1), (1, 0, 1),
         (3, 0, 0), (0, 3, 0), (0, 0, 3), (2, 1, 0), (2, 0, 1), (1, 2, 0), (0, 2, 1), (1, 0, 2), (0, 1, 2), (1, 1, 1),
         (4, 0, 0), (0, 4, 0), (0, 0, 4), (3, 1, 0), (3, 0, 1), (1, 3, 0), (0, 3, 1), (1, 0, 3), (0, 1, 3), (2, 2, 0),
         (0, 2, 2), (2, 0, 2), (2, 1, 1), (1, 2, 1), (1, 1, 2)]


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_coarse_moments(fm, S, degree, prec):
    """The moment kernel of the coarse level (per-thread sums in the PCG's
    precision, DPP wave sums, FP64 block sums) against a float64 sum over the
    same FP32-rounded scaled coordinates, on a jittered box of odd size; the
    constrained components contribute nothing.  A wrong lane pattern in the
    wave sums drops or doubles lanes' partials, which this bound catches."""
    cfg = cube(fm, dims=(9, 7, 11))
    fm.configs.load(S, cfg)
    nn = cfg["X"].size // 3
    X = cfg["X"].reshape(3, nn)
    xs = np.empty((3, nn))
    for d in range(3):
        lo, hi = X[d].min(), X[d].max()
        xs[d] = ((X[d] - 0.5 * (lo + hi)) * (2.0 / (hi - lo))).astype(np.float32)
    r = np.random.default_rng(degree).standard_normal(3 * nn)
    mask = np.zeros(3 * nn, bool)
    mask[cfg["dir_dofs"]] = True
    rm = np.where(mask, 0.0, r).reshape(nn, 3)
    nm = (4, 10, 20, 35)[degree - 1]
    phi = np.stack([xs[0] ** a * xs[1] ** b * xs[2] ** c for a, b, c in MONOS[:nm]])  # nm x nn
    want = (phi @ rm).T.reshape(-1)  # index c nm + k
    scale = np.abs(phi) @ np.abs(rm)
    p = fm.FM_FP64 if prec == "fp64" else fm.FM_MIXED
    got = S.coarse_moments(r, degree, p)
    tol = 1e-13 if prec == "fp64" else 2e-6
    assert np.all(np.abs(got - want) <= tol * scale.T.reshape(-1) + 1e-300), (prec, degree)
    # the constrained components: r there changes nothing
    r2 = r.copy()
    r2[cfg["dir_dofs"]] += 1.0
    assert np.array_equal(S.coarse_moments(r2, degree, p), got)


def test_fused_polynomial_gram_matches_generic(tmp_path):
    """The fused pass for the polynomial fields' Gram rows (no A Z formed)
    against the generic path (every basis column's product, one dense Gram
    matrix) over eight load steps with the bench's inner-solve options: the
    same Newton and PCG iteration counts in FP64 (within two PCG iterations
    per step in FP32, whose Gram entries round differently) and the same
    final states to the solves' accuracy."""
    import subprocess
    import sys

    here = os.path.dirname(os.path.abspath(__file__))
    outs = {}
    for mode in ("fused", "generic"):
        env = dict(os.environ)
        env.pop("FM_GALERKIN", None)
        if mode == "generic":
            env["FM_GALERKIN"] = "generic"
        f = str(tmp_path / f"{mode}.npz")
        p = subprocess.run([sys.executable, os.path.join(here, "galerkin_check.py"), f], env=env,
                           capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stdout + p.stderr
        outs[mode] = np.load(f)
    a, b = outs["fused"], outs["generic"]
    assert np.array_equal(a["fp64_its"], b["fp64_its"]), (a["fp64_its"], b["fp64_its"])
    assert np.array_equal(a["mixed_its"][:, 0], b["mixed_its"][:, 0])
    assert np.abs(a["mixed_its"][:, 1] - b["mixed_its"][:, 1]).max() <= 2, (a["mixed_its"], b["mixed_its"])
    for name in ("mixed", "fp64"):
        assert rel(a[name + "_u"], b[name + "_u"]) < 1e-6, name


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_ritz_vectors_in_the_start(fm, S, O, prec):
    """With ritz = 6, every solve keeps its Lanczos vectors and the next
    step's first solve also projects onto the Ritz vectors of the slowest
    modes: every step still reaches ||R|| <= 1e-8 ||R_0||, the final state
    matches the FP64 oracle's, the path repeats bit for bit after a reset,
    and the Lanczos coefficients of a solve give a Lanczos matrix whose
    eigenvalues lie in (0, ~6] (the preconditioned tangent is SPD).

    The path runs three times from the unloaded state: its first 3 steps,
    then twice in full.  The first Newton solve after a reset projects onto
    the predictor solve's Ritz vectors with no recycled correction beside
    them, and each pass leaves the device's Gram buffers in another state,
    so the passes agree only if that start forms every entry it reads."""
    cfg = cube(fm, n_steps=8, dims=(12, 12, 12))
    fm.configs.load(S, cfg)
    P = oracle(O, cfg, pcg_tol=1e-12, pcg_max=4000)
    uo = np.zeros(cfg["X"].size)
    for s in range(1, 9):
        P.step(s, uo)
    kw = dict(prec=fm.FM_MIXED) if prec == "mixed" else dict(prec=fm.FM_FP64, inexact=1)
    c = fm.configs.newton_cfg(cfg, lift=2, inner_tol=1e-3, inner_tol_max=0.1, recycle=6, coarse_degree=4,
                              coarse_additive=1, lag_precond=1, factor_steps=6, extrap_order=3, ritz=4, **kw)
    runs = []
    for last in (3, 8, 8):
        S.set_u(None)
        stats = []
        for s in range(1, last + 1):
            rc, st = S.newton_step(s, c)
            assert st["converged"] and st["rel_res"] <= 1e-8 and st["pcg_cap_hits"] == 0, (s, st)
            stats.append((st["newton_iters"], st["pcg_iters"]))
            a, b = S.pcg_history()
            if a.size >= 3:
                d = 1.0 / a
                d[1:] += b[:-1] / a[:-1]
                e = -np.sqrt(np.maximum(b[:-1], 0)) / a[:-1]
                w = np.linalg.eigvalsh(np.diag(d) + np.diag(e, 1) + np.diag(e, -1))
                assert w.min() > 0 and w.max() < 10, (s, w.min(), w.max())
        runs.append((S.get_u(), stats))
    assert rel(runs[1][0], uo) < 1e-6
    assert runs[1][1][:3] == runs[0][1], (runs[0][1], runs[1][1])
    assert np.array_equal(runs[1][0], runs[2][0]) and runs[1][1] == runs[2][1], (runs[1][1], runs[2][1])


def test_fp32_preconditioned_beam_with_coarse_level(fm, S, O):
    """Configuration 5 form at 48x6x6 (FP64 Krylov vectors and SpMV, the FP64
    ILU(0) factor rounded to FP32 as preconditioner, energy rule): with the
    configuration's inner-solve options (extrapolated start, Galerkin start
    over recycled corrections and the degree-4 polynomial fields, the same
    fields as additive coarse level, whose r.(Z c) terms enter the flexible
    beta) every step reaches ||R|| <= 1e-8 ||R_0|| with fewer PCG iterations
    than the plain FP32-preconditioned solves, and both end at the FP64
    oracle's state."""
    cfg = fm.configs.make(5, dims=(48, 6, 6))
    cfg["n_steps"] = 40
    assert cfg["prec"] == fm.FM_MIXED_PC and cfg["inner_opts"]["coarse_additive"] == 1
    fm.configs.load(S, cfg)
    steps = range(1, 4)
    P = oracle(O, cfg, pcg_tol=1e-12, pcg_max=6000, ls_mode=1, max_newton=60)
    uo = np.zeros(cfg["X"].size)
    for s in steps:
        P.step(s, uo)
    # the plain form: the configuration without its inner-solve options
    plain = fm.configs.newton_cfg({**cfg, "inner_opts": {}}, inner_tol=1e-2)
    pcg = {}
    for name, c in (("coarse", fm.configs.newton_cfg(cfg)), ("plain", plain)):
        S.set_u(None)
        pcg[name] = 0
        for s in steps:
            rc, st = S.newton_step(s, c)
            assert st["converged"] and st["rel_res"] <= 1e-8 and st["pcg_cap_hits"] == 0, (name, s, st)
            pcg[name] += st["pcg_iters"]
        assert rel(S.get_u(), uo) < 1e-6, name
    assert pcg["coarse"] < 0.5 * pcg["plain"], pcg
