"""Parity of the HIP path (through the C ABI) with the FP64 CPU oracle.

Integer outputs (CSR pattern, colouring) must be bit-exact.  Floating-point
tolerances, relative to the largest magnitude of the compared quantity:
  FP64 kernels            <= 1e-12 (summation order differs from the oracle)
  FP32 tangent / SpMV     <= 2e-6
  FP32 ILU(0) factors      <= 1e-3 (pivot-error growth over the factorization)
  Newton displacement     <= 1e-9 (FP64 path), <= 1e-7 (mixed path, rtol 1e-8)
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import golden

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


def small_problem(fm, k=3, dims=(4, 3, 5)):
    cfg = fm.configs.make(k, dims=dims)
    return cfg


def oracle_problem(O, cfg, **kw):
    return O.NewtonProblem(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], cfg["dir_dofs"], cfg["dir_vals"],
                           cfg["f_ext"], cfg["n_steps"], **{**{k: cfg[k] for k in ("ls_mode", "max_newton", "pcg_max") if k in cfg}, **kw})


# ---------------------------------------------------------------------------- pattern / colouring
@pytest.mark.parametrize("k,dims", [(1, (8, 8, 8)), (2, (6, 5, 7)), (5, (16, 3, 2))])
def test_pattern_and_coloring_bit_exact(fm, S, O, k, dims):
    cfg = fm.configs.make(k, dims=dims)
    fm.configs.load(S, cfg)
    rp, ci = S.pattern()
    nn = cfg["X"].size // 3
    rpo, cio = O.pattern(nn, cfg["conn"])
    assert np.array_equal(rp, rpo) and np.array_equal(ci, cio)
    col, nc = S.colors()
    colo, nco = O.color(nn, cfg["conn"])
    assert nc == nco and np.array_equal(col, colo)


def test_coloring_matches_compiled_reference(fm, S):
    g = golden("mesh_pattern_color.npz")
    X, conn = fm.mesh_box(8, 8, 8)
    S.set_mesh(X, conn)
    col, nc = S.colors()
    assert nc == int(g["ref_ncolors_kuhn_8"][0]) and np.array_equal(col, g["ref_colors_kuhn_8"])
    rp, _ = S.pattern()
    assert np.array_equal(rp, g["ref_rowptr_kuhn_8"])


# ---------------------------------------------------------------------------- assembly
@pytest.mark.parametrize("k", [2, 3])
def test_assembly_matches_oracle(fm, S, O, k):
    cfg = small_problem(fm, k)
    fm.configs.load(S, cfg)
    rng = np.random.default_rng(k)
    nn = cfg["X"].size // 3
    u = rng.uniform(-0.02, 0.02, 3 * nn)
    rp, ci = O.pattern(nn, cfg["conn"])
    colors, nc = O.color(nn, cfg["conn"])
    Ko, fo = O.assemble(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], colors, nc, u, rp, ci)
    f64, K64 = S.assemble(u, fm.FM_FP64)
    assert rel(f64, fo) < 1e-12 and rel(K64, Ko) < 1e-12
    f32, K32 = S.assemble(u, fm.FM_MIXED)
    assert rel(f32, fo) < 1e-12  # the force is FP64 in both paths
    assert rel(K32, Ko) < 2e-6
    R, nr = S.residual(u, 0.7)
    Ro = -fo.copy()
    Ro[cfg["dir_dofs"]] = 0.0
    assert rel(R, Ro) < 1e-12 and abs(nr - np.linalg.norm(Ro)) < 1e-12 * np.linalg.norm(Ro)


def test_assembly_is_bitwise_repeatable(fm, S):
    cfg = small_problem(fm, 2, (5, 4, 4))
    fm.configs.load(S, cfg)
    u = np.random.default_rng(1).uniform(-0.02, 0.02, cfg["X"].size)
    f1, K1 = S.assemble(u)
    f2, K2 = S.assemble(u)
    assert np.array_equal(f1, f2) and np.array_equal(K1, K2)


def test_use_tangent_eliminates_constrained_rows_and_columns(fm, S, O):
    """fm_use_tangent(eliminate=1) on the assembled tangent: zero rows and
    columns with a unit diagonal at the constrained DOFs
    (FEM/nonlinear_solver.hpp:388-396), in both precisions, seen through the
    SpMV of the matrix it leaves current."""
    cfg = small_problem(fm, 2, (5, 4, 6))
    fm.configs.load(S, cfg)
    nn = cfg["X"].size // 3
    rng = np.random.default_rng(7)
    u = rng.uniform(-0.02, 0.02, 3 * nn)
    rp, ci = O.pattern(nn, cfg["conn"])
    colors, nc = O.color(nn, cfg["conn"])
    Ko, _ = O.assemble(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], colors, nc, u, rp, ci)
    isd = np.zeros(3 * nn, dtype=np.uint8)
    isd[cfg["dir_dofs"]] = 1
    O.eliminate(rp, ci, Ko, isd)
    x = rng.standard_normal(3 * nn)
    yo = O.spmv(rp, ci, Ko, x)
    d = cfg["dir_dofs"]
    for prec, tol, xd in ((fm.FM_FP64, 1e-12, x[d]), (fm.FM_MIXED, 2e-6, x[d].astype(np.float32).astype(np.float64))):
        S.assemble(u, prec)
        S.use_tangent(True)
        y = S.spmv(x, prec)
        assert rel(y, yo) < tol
        assert np.array_equal(y[d], xd)  # unit diagonal, nothing else in the row


# ---------------------------------------------------------------------------- SpMV / ILU(0) / PCG
def poisson(g):
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as MG

    return MG.poisson(g, False)


def elastic_matrix(O, fm, dims=(4, 4, 4), seed=3, eliminate=True):
    cfg = fm.configs.make(2, dims=dims)
    nn = cfg["X"].size // 3
    rp, ci = O.pattern(nn, cfg["conn"])
    colors, nc = O.color(nn, cfg["conn"])
    u = np.random.default_rng(seed).uniform(-0.02, 0.02, 3 * nn)
    K, _ = O.assemble(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], colors, nc, u, rp, ci)
    if eliminate:
        isd = np.zeros(3 * nn, dtype=np.uint8)
        isd[cfg["dir_dofs"]] = 1
        O.eliminate(rp, ci, K, isd)
    return rp, ci, K


def sorted_csr(rp, ci, va):
    perm = np.concatenate([rp[r] + np.argsort(ci[rp[r]:rp[r + 1]], kind="stable") for r in range(rp.size - 1)])
    return ci[perm], va[perm], perm


@pytest.mark.parametrize("which", ["poisson", "elastic"])
def test_spmv_ilu0_match_oracle(fm, S, O, which):
    if which == "poisson":
        rp, ci, va, b, xs = poisson(24)
    else:
        rp, ci, va = elastic_matrix(O, fm)
    S.set_matrix(rp, ci, va)
    si = S.sweep_info()
    assert si["sweep"] == ("level" if which == "poisson" else "chain")
    assert si["spmv"] == ("csr" if which == "poisson" else "node")
    n = rp.size - 1
    x = np.random.default_rng(2).standard_normal(n)
    cs, vs, perm = sorted_csr(rp, ci, va)
    yo = O.spmv(rp, cs, vs, x)
    assert rel(S.spmv(x), yo) < 1e-13
    assert rel(S.spmv(x, fm.FM_MIXED), yo) < 2e-6
    LUo, diag = O.ilu0(rp, cs, vs)
    LU = S.ilu0_factor()
    LUo_user = np.empty_like(LUo)
    LUo_user[perm] = LUo
    assert rel(LU, LUo_user) < 1e-12
    zo = O.ilu0_apply(rp, cs, LUo, diag, x)
    assert rel(S.ilu0_apply(x), zo) < 1e-12
    LU32 = S.ilu0_factor(fm.FM_MIXED)
    assert rel(LU32, LUo_user) < 1e-3
    assert rel(S.ilu0_apply(x, fm.FM_MIXED), zo) < 1e-3


def dense_coupled_blocks(nn=40, seed=5):
    """Node-blocked (3x3) SPD matrix: a block chain plus node 0 coupled to nodes
    1..20 (21 upper blocks) and node nn-1 coupled to nodes 0..19 (more than 8
    lower neighbours), so the factorization mixes nodes inside and outside the
    limits of the LDS-staged path."""
    rng = np.random.default_rng(seed)
    nb = [set([i]) for i in range(nn)]
    for i in range(nn - 1):
        nb[i].add(i + 1)
        nb[i + 1].add(i)
    for j in range(1, 21):
        nb[0].add(j)
        nb[j].add(0)
    for j in range(0, 20):
        nb[nn - 1].add(j)
        nb[j].add(nn - 1)
    blocks = {}
    for i in range(nn):
        for j in nb[i]:
            if j > i:
                blocks[(i, j)] = rng.uniform(-1, 1, (3, 3))
                blocks[(j, i)] = blocks[(i, j)].T
    for i in range(nn):
        g = rng.uniform(-1, 1, (3, 3))
        blocks[(i, i)] = g @ g.T + 3.0 * (len(nb[i]) + 1) * np.eye(3)
    rp, ci, va = [0], [], []
    for i in range(nn):
        cols = sorted(nb[i])
        for d in range(3):
            for j in cols:
                ci.extend(3 * j + e for e in range(3))
                va.extend(blocks[(i, j)][d])
            rp.append(len(ci))
    return np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va)


@pytest.mark.parametrize("which", ["elastic", "dense"])
def test_factor_staged_matches_unstaged(fm, S, O, which, monkeypatch):
    """The LDS-staged 3x3-node factorization performs the unstaged code's
    arithmetic in the same order: identical bits in FP64 and FP32, and the
    oracle factor to the FP64 tolerance.  The dense case has nodes beyond the
    staging limits (21 upper blocks; 20 lower neighbours), which the same
    launch factorizes unstaged."""
    rp, ci, va = elastic_matrix(O, fm, dims=(5, 4, 4)) if which == "elastic" else dense_coupled_blocks()
    S.set_matrix(rp, ci, va)
    assert S.sweep_info()["spmv"] == "node"
    cs, vs, perm = sorted_csr(rp, ci, va)
    LUo, _ = O.ilu0(rp, cs, vs)
    LUo_user = np.empty_like(LUo)
    LUo_user[perm] = LUo
    got = {}
    for staged in ("1", "0"):
        monkeypatch.setenv("FM_FACTOR_STAGED", staged)
        got[staged] = (S.ilu0_factor(), S.ilu0_factor(fm.FM_MIXED))
    monkeypatch.delenv("FM_FACTOR_STAGED")
    assert rel(got["1"][0], LUo_user) < 1e-12
    assert rel(got["1"][1], LUo_user) < 1e-3
    for k in range(2):
        assert np.array_equal(got["1"][k].view(np.uint64), got["0"][k].view(np.uint64))


@pytest.mark.parametrize("dims", [(9, 7, 6), (24, 3, 3), (5, 11, 13)])
def test_chain_sweeps_match_oracle(fm, S, O, dims):
    # several bundles per direction, partial bundles at the edges, a slender
    # beam; clamped (eliminated) tangents, as the Newton solve factorizes them
    rp, ci, va = elastic_matrix(O, fm, dims=dims)
    S.set_matrix(rp, ci, va)
    info = S.sweep_info()
    assert info["sweep"] == "chain" and info["chains"] == (dims[1] + 1) * (dims[2] + 1)
    n = rp.size - 1
    LUo, diag = O.ilu0(rp, ci, va)
    # the FP32 sweeps against the oracle apply of the GPU's own FP32 factor:
    # what is left is the FP32 arithmetic of the two triangular solves
    LU32 = S.ilu0_factor(fm.FM_MIXED)
    dpos = np.array([rp[i] + np.searchsorted(ci[rp[i]:rp[i + 1]], i) for i in range(n)])
    for seed in range(3):
        x = np.random.default_rng(seed).standard_normal(n)
        zo = O.ilu0_apply(rp, ci, LUo, diag, x)
        assert rel(S.ilu0_apply(x), zo) < 1e-12
        x32 = x.astype(np.float32).astype(np.float64)
        z32o = O.ilu0_apply(rp, ci, LU32, LU32[dpos], x32)
        z32 = S.ilu0_apply(x, fm.FM_MIXED)
        print(f"seed {seed}: fp32 sweeps vs own-factor oracle {rel(z32, z32o):.2e}, vs fp64 {rel(z32, zo):.2e}")
        assert rel(z32, z32o) < 1e-4


def test_level_sweeps_still_match_oracle(fm):
    # the level-scheduled sweeps (FM_SWEEP=level) on a 3-row node-blocked matrix
    here = os.path.dirname(os.path.abspath(__file__))
    env = dict(os.environ, FM_SWEEP="level")
    p = subprocess.run([sys.executable, os.path.join(here, "sweep_check.py"), "level"], env=env,
                       capture_output=True, text=True, timeout=300)
    print(p.stdout, p.stderr)
    assert p.returncode == 0 and "SWEEP OK level" in p.stdout


def test_chain_sweep_kernels_same_bits(tmp_path):
    """The one-wave and the two-wave (loader + solver) chain-sweep kernels do
    the same arithmetic in the same order: identical ILU(0) applies and PCG
    solves (whose backward sweeps carry the fused r.y partials), FP64 and FP32."""
    here = os.path.dirname(os.path.abspath(__file__))
    res = {}
    for v in ("0", "1"):
        out = tmp_path / f"loader{v}.npz"
        env = dict(os.environ, FM_SWEEP_LOADER=v)
        env.pop("FM_SWEEP", None)
        p = subprocess.run([sys.executable, os.path.join(here, "sweep_check.py"), "chain", str(out)], env=env,
                           capture_output=True, text=True, timeout=300)
        print(p.stdout, p.stderr)
        assert p.returncode == 0 and "SWEEP OK chain" in p.stdout
        res[v] = dict(np.load(out))
    assert res["0"].keys() == res["1"].keys() and len(res["0"]) == 15
    assert res["0"]["7x6x5_kind"][0] and res["0"]["20x3x3_kind"][0]
    for k in res["0"]:
        assert np.array_equal(res["0"][k], res["1"][k]), k


@pytest.mark.parametrize("gn", [16, 32, 48])
def test_pcg_matches_reference_golden(fm, S, O, gn):
    g = golden("pcg.npz")
    rp, ci, va, b, xs = poisson(gn)
    S.set_matrix(rp, ci, va)
    x, its, relres = S.pcg(b, tol=1e-10, max_it=4000)
    its_ref = int(g[f"poisson{gn}_refcols_iters"][0])
    assert abs(its - its_ref) <= 1
    assert rel(x, g[f"poisson{gn}_refcols_x"]) < 1e-9
    assert relres < 1e-10


def test_pcg_elastic_matches_oracle(fm, S, O):
    rp, ci, va = elastic_matrix(O, fm, dims=(6, 5, 5))
    b = np.random.default_rng(9).standard_normal(rp.size - 1)
    S.set_matrix(rp, ci, va)
    xo, itso, _ = O.pcg(rp, ci, va, b, tol=1e-10, max_iter=500)
    x, its, relres = S.pcg(b, tol=1e-10, max_it=500)
    assert abs(its - itso) <= 1
    assert rel(x, xo) < 1e-9
    x32, its32, _ = S.pcg(b, tol=1e-5, max_it=500, prec=fm.FM_MIXED)
    assert rel(x32, xo) < 1e-3


@pytest.mark.parametrize("dims", [(6, 5, 5), (20, 3, 3)])
def test_pcg_fp32_preconditioner_matches_oracle(fm, S, O, dims):
    # FP64 Krylov with the FP64 factor rounded to FP32 and applied in FP32:
    # the solution is the FP64 one.  On a well-conditioned tangent the
    # iteration count is the FP64 one.  On an ill-conditioned tangent (the
    # slender 20x3x3 beam) the FP32 application makes the preconditioner
    # inexact, and the count then depends on the rounding of the other
    # kernels (185 to 681 observed for summation orders of the FP64 SpMV), so
    # it is only bounded there
    rp, ci, va = elastic_matrix(O, fm, dims=dims)
    b = np.random.default_rng(9).standard_normal(rp.size - 1)
    S.set_matrix(rp, ci, va)
    xo, itso, _ = O.pcg(rp, ci, va, b, tol=1e-10, max_iter=2000)
    x, its, relres = S.pcg(b, tol=1e-10, max_it=2000, prec=fm.FM_MIXED_PC)
    print(f"{dims}: FP32-preconditioned PCG {its} iterations, FP64 oracle {itso}")
    assert relres < 1e-10 and its <= (1 if dims == (6, 5, 5) else 10) * itso + 1
    assert rel(x, xo) < 1e-8


def test_newton_fp32_preconditioner_matches_fp64_oracle(fm, S, O):
    cfg = small_problem(fm, k=3, dims=(6, 6, 6))
    fm.configs.load(S, cfg)
    uo, sto = run_oracle(O, cfg, range(1, 3))
    c = fm.configs.newton_cfg(cfg, prec=fm.FM_MIXED_PC)
    for s in range(1, 3):
        rc, st = S.newton_step(s, c)
        assert st["converged"]
    assert rel(S.get_u(), uo) < 1e-7


def test_pcg_edge_cases(fm, S):
    rp = np.array([0, 1, 2, 3], np.uint32)
    ci = np.array([0, 1, 2], np.uint32)
    S.set_matrix(rp, ci, np.array([2.0, 4.0, 8.0]))
    x, its, _ = S.pcg(np.zeros(3), tol=1e-10)  # b = 0: trivially converged
    assert its == 0 and not x.any()
    x, its, _ = S.pcg(np.array([2.0, 4.0, 8.0]), tol=1e-10)
    assert its == 1 and np.allclose(x, 1.0, rtol=0, atol=1e-15)
    # zero pivot: the reference perturbation u_ii += 1e-8
    rp = np.array([0, 2, 4], np.uint32)
    ci = np.array([0, 1, 0, 1], np.uint32)
    S.set_matrix(rp, ci, np.ones(4))
    LU = S.ilu0_factor()
    assert LU[3] == 1e-8 and LU[0] == 1.0


# ---------------------------------------------------------------------------- Newton
def run_oracle(O, cfg, steps, **kw):
    P = oracle_problem(O, cfg, newton_rtol=cfg["newton_rtol"], pcg_tol=cfg["pcg_tol"], **kw)
    u = np.zeros(cfg["X"].size)
    stats = []
    for s in steps:
        rc, st = P.step(s, u)
        assert rc == 0
        stats.append(st)
    return u, stats


def test_newton_config1_fp64_matches_oracle(fm, S, O):
    cfg = fm.configs.make(1)
    fm.configs.load(S, cfg)
    uo, sto = run_oracle(O, cfg, range(1, 11))
    c = fm.configs.newton_cfg(cfg)
    u = np.zeros(cfg["X"].size)
    for s in range(1, 11):
        rc, st = S.newton_step(s, c, u)
        assert st["converged"] and st["rel_res"] <= 1e-10
        assert st["newton_iters"] == sto[s - 1]["newton_iters"]
    assert rel(u, uo) < 1e-9


def test_newton_mixed_matches_fp64_oracle(fm, S, O):
    cfg = fm.configs.make(3, dims=(6, 6, 6))
    cfg["n_steps"] = 4
    fm.configs.load(S, cfg)
    uo, _ = run_oracle(O, cfg, range(1, 5))
    c = fm.configs.newton_cfg(cfg, inner_tol=1e-4)
    for s in range(1, 5):
        rc, st = S.newton_step(s, c)
        assert st["converged"] and st["rel_res"] <= 1e-8
    u = S.get_u()
    assert rel(u, uo) < 1e-7
    # the FP64 residual of the mixed solution, evaluated by the oracle
    nn = cfg["X"].size // 3
    rp, ci = O.pattern(nn, cfg["conn"])
    colors, nc = O.color(nn, cfg["conn"])
    _, f = O.assemble(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], colors, nc, u, rp, ci, want_K=False)
    f[cfg["dir_dofs"]] = 0.0
    _, f0 = O.assemble(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], colors, nc, uo, rp, ci, want_K=False)
    assert np.linalg.norm(f) <= 1e-6 * np.abs(f0).max()


def test_newton_mixed_extrapolating_predictor(fm, S, O):
    """lift = 2: from the second step on, the load step starts from the
    extrapolation of the previous converged states (linear, then quadratic; no
    predictor solve), and from the quadratic start later Newton iterations may
    keep the first tangent and factor; the solution still matches the FP64
    oracle, and any fm_set_* restarts the history."""
    cfg = fm.configs.make(3, dims=(6, 6, 6))
    cfg["n_steps"] = 5
    fm.configs.load(S, cfg)
    uo, _ = run_oracle(O, cfg, range(1, 6))
    c = fm.configs.newton_cfg(cfg, inner_tol=1e-4, lift=2)
    for s in range(1, 6):
        rc, st = S.newton_step(s, c)
        assert st["converged"] and st["rel_res"] <= 1e-8
        # step 1 knows only the zero state: it takes the predictor solve;
        # step 2 starts from the secant, later steps from the quadratic
        if s <= 2:
            assert st["factorizations"] == st["newton_iters"] + (1 if s == 1 else 0), (s, st)
        else:
            assert 1 <= st["factorizations"] == st["assemblies"] <= st["newton_iters"], (s, st)
        if s == 4:
            u4 = S.get_u()
    u5 = S.get_u()
    assert rel(u5, uo) < 1e-7
    S.set_u(u4)  # history restarts: step 5 again takes the predictor solve
    rc, st = S.newton_step(5, c)
    assert st["converged"] and st["factorizations"] == st["newton_iters"] + 1
    assert rel(S.get_u(), u5) < 1e-7


@pytest.mark.parametrize("ls_mode", [0, 1])
def test_newton_traction_beam_fp64(fm, S, O, ls_mode):
    # traction-driven (no Dirichlet increment, so no lift): the first two of 40
    # steps, with the residual-norm rule (0) and the energy rule (1)
    cfg = fm.configs.make(5, dims=(16, 2, 2))
    cfg["prec"] = fm.FM_FP64
    cfg["ls_mode"] = ls_mode
    fm.configs.load(S, cfg)
    uo, sto = run_oracle(O, cfg, range(1, 3))
    c = fm.configs.newton_cfg(cfg)
    for s in range(1, 3):
        rc, st = S.newton_step(s, c)
        assert st["converged"] and st["newton_iters"] == sto[s - 1]["newton_iters"]
    if ls_mode == 1:
        assert max(t["newton_iters"] for t in sto) <= 5
    assert rel(S.get_u(), uo) < 1e-9


def test_newton_traction_beam_mixed_energy_rule(fm, S, O):
    # the beam's own setting: FP64 Krylov with the FP32 ILU(0) preconditioner
    # and the energy rule
    cfg = fm.configs.make(5, dims=(32, 4, 4))
    assert cfg["ls_mode"] == 1 and cfg["prec"] == fm.FM_MIXED_PC
    fm.configs.load(S, cfg)
    c = fm.configs.newton_cfg(cfg)
    for s in range(1, 3):
        rc, st = S.newton_step(s, c)
        assert st["converged"] and st["newton_iters"] <= 8
    cfg["prec"] = fm.FM_FP64
    uo, _ = run_oracle(O, cfg, range(1, 3), pcg_max=3000)
    assert rel(S.get_u(), uo) < 1e-6


# ---------------------------------------------------------------------------- adjoint / rotation / checkpoint
def test_adjoint_matches_oracle_and_fd(fm, S, O):
    cfg = fm.configs.make(4, dims=(3, 3, 3))
    cfg["n_steps"] = 2
    fm.configs.load(S, cfg)
    c = fm.configs.newton_cfg(cfg, pcg_max=2000)
    for s in (1, 2):
        S.newton_step(s, c)
    u = S.get_u()
    nn = cfg["X"].size // 3
    ustar = np.random.default_rng(4).uniform(-0.01, 0.01, 3 * nn)
    top = cfg["loss_nodes"]
    dLdu = np.zeros(3 * nn)
    dLdu.reshape(nn, 3)[top] = (u - ustar).reshape(nn, 3)[top]
    dmu, dlam, dload, its = S.adjoint(dLdu)
    P = oracle_problem(O, cfg, newton_rtol=1e-13, pcg_tol=1e-13, pcg_max=2000)
    uo = np.zeros(3 * nn)
    for s in (1, 2):
        P.step(s, uo)
    dmo, dlo, dldo, _, _ = P.adjoint(uo, dLdu)
    assert rel(dmu, dmo) < 1e-8 and rel(dlam, dlo) < 1e-8 and abs(dload - dldo) <= 1e-8 * abs(dldo)


def test_rotations_match_oracle(fm, S, O):
    rng = np.random.default_rng(5)
    Fs = []
    for ang in (1e-9, 1e-3, 0.4, 1.2, 2.5, np.pi - 1e-4):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        Rt = O.rot_exp(ang * n)
        A = rng.standard_normal((3, 3))
        Fs.append(Rt @ (np.eye(3) + 0.2 * A @ A.T))
    R, w, Q = S.rot_logexp(np.array(Fs), alpha=0.5)
    for i, F in enumerate(Fs):
        Ro, _ = O.polar(F)
        wo = O.rot_log(Ro)
        assert np.abs(R[i] - Ro).max() < 1e-14
        assert np.abs(w[i] - wo).max() < 1e-13
        assert np.abs(Q[i] - O.rot_exp(0.5 * wo)).max() < 1e-14


def test_element_fields_match_oracle(fm, S, O):
    """Per-element fields (det F, Cauchy stress, von Mises, pressure, largest
    principal stress, energy, maximum stretch, F) against the FP64 oracle, for
    a given u and for the device state; a rigid rotation gives zero stress."""
    cfg = fm.configs.make(2, dims=(5, 4, 3))
    fm.configs.load(S, cfg)
    rng = np.random.default_rng(8)
    u = rng.uniform(-0.04, 0.04, cfg["X"].size)
    fo = O.element_fields(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], u)
    fg = S.element_fields(u)
    for k in fo:
        assert rel(fg[k], fo[k]) < 1e-12, k
    S.set_u(u)
    fs = S.element_fields(None, fields=("von_mises", "F"))
    assert set(fs) == {"von_mises", "F"}
    assert rel(fs["von_mises"], fo["von_mises"]) < 1e-12 and rel(fs["F"], fo["F"]) < 1e-12
    # rigid rotation u = (Q - I) X: F = Q, det F = 1, stress and energy zero, stretch one
    th = 0.9
    Q = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
    X = cfg["X"].reshape(3, -1)
    ur = ((Q - np.eye(3)) @ X).T.ravel()
    fr = S.element_fields(ur)
    mu_max = cfg["mu"].max()
    assert np.abs(fr["det_F"] - 1.0).max() < 1e-13
    assert np.abs(fr["cauchy"]).max() < 1e-13 * mu_max and np.abs(fr["von_mises"]).max() < 1e-12 * mu_max
    assert np.abs(fr["energy"]).max() < 1e-13 * mu_max
    assert np.abs(fr["max_stretch"] - 1.0).max() < 1e-13
    assert np.abs(fr["F"] - Q.ravel()[:, None]).max() < 1e-13


def test_zz_estimate_matches_oracle(fm, S, O):
    """ZZ recovered-gradient estimate: node averages, SPR-recovered values,
    element indicators and the global estimate against the FP64 oracle (the
    patches are summed in another order: rounding-level differences, amplified
    by the patch normal equations); the device state; and a context with a
    mesh but no material."""
    cfg = fm.configs.make(2, dims=(5, 4, 3))
    fm.configs.load(S, cfg)
    Xn = cfg["X"].reshape(3, -1)
    rng = np.random.default_rng(9)
    u = np.stack([0.05 * Xn[0] ** 2, 0.03 * Xn[0] * Xn[1], -0.02 * Xn[2] ** 2]).T.ravel()
    u += rng.uniform(-1e-3, 1e-3, u.size)
    go, eo, ro, co = O.zz_estimate(cfg["X"], cfg["conn"], u)
    gg, eg, rg, cg = S.zz_estimate(u)
    scale = np.abs(ro).max()
    assert np.abs(rg - ro).max() < 1e-13 * scale
    assert np.abs(cg - co).max() < 1e-10 * scale
    assert np.abs(eg - eo).max() < 1e-9 * eo.max() and abs(gg - go) < 1e-9 * go
    S.set_u(u)
    gs, es, _, _ = S.zz_estimate()
    assert gs == gg and np.array_equal(es, eg)
    T = fm.Solver(0)
    try:
        T.set_mesh(cfg["X"], cfg["conn"])
        gt, et, _, _ = T.zz_estimate(u)
        assert gt == gg and np.array_equal(et, eg)
    finally:
        T.close()


def test_one_context_across_meshes_and_matrices(fm, O):
    """One context: mesh A (pattern, colours, assembly), an external CSR
    matrix through PCG, then mesh B through Newton -- the sequence a host
    engine drives through the C++ shim."""
    S = fm.Solver(0)
    try:
        a = fm.configs.make(2, dims=(3, 3, 3))
        fm.configs.load(S, a)
        S.pattern()
        S.colors()
        S.assemble(np.zeros(a["X"].size))
        rp, ci, va, b, xs = poisson(16)
        S.set_matrix(rp, ci, va)
        x, its, _ = S.pcg(b, tol=1e-10, max_it=4000)
        assert its == int(golden("pcg.npz")["poisson16_refcols_iters"][0])
        c = fm.configs.make(1, dims=(3, 3, 3))
        fm.configs.load(S, c)
        rc, st = S.newton_step(1, fm.configs.newton_cfg(c))
        assert st["converged"]
        uo, _ = run_oracle(O, c, [1])
        assert rel(S.get_u(), uo) < 1e-9
    finally:
        S.close()
