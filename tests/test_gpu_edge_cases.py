"""Edge cases of the element kernel and of the assembly paths, and the
configuration workloads not covered elsewhere, through the C ABI against the
FP64 CPU oracle.

* J < 0.01: ln J is clamped at ln 0.01 and the tangent drops the lambda' term
  (FEM/constitutive_models.hpp:96; the oracle's clamp is pinned to the compiled
  compute_pk1_stress in tests/golden/pk1.npz and its tangent by FD in
  tests/test_oracle.py).  Inverted elements (det F < 0) take J = |det F|.
* zero-volume elements contribute nothing (FEM/adaptive_fem_engine.hpp:545-549,
  1016, 1169).
* the colour-launch assembly (FM_ASM=color, and the automatic fallback for a
  node with more incident elements or column blocks than the node-gather
  kernel holds) against the oracle and against the node-gather path.
* a compression load step with the predictor solve (configuration 2 form) and
  the adjoint of the traction beam (configuration 5 form) at reduced size.

Tolerances as in test_gpu_parity.py: FP64 kernels <= 1e-12, FP32 tangent
<= 2e-6 (relative to the largest magnitude), Newton displacement <= 1e-9,
adjoint vs oracle <= 1e-8 and vs central FD (h = 1e-6) <= 1e-5.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


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


def det_F(X, conn, u):
    """det F of every element (numpy, from the P1 geometry)."""
    nn, ne = X.size // 3, conn.size // 4
    P = X.reshape(3, nn).T
    x = P + u.reshape(nn, 3)
    c = conn.reshape(4, ne).T
    J0 = np.stack([P[c[:, a]] - P[c[:, 0]] for a in (1, 2, 3)], axis=2)
    J1 = np.stack([x[c[:, a]] - x[c[:, 0]] for a in (1, 2, 3)], axis=2)
    return np.linalg.det(J1) / np.linalg.det(J0)


def oracle_assembly(O, X, conn, mu, lam, u):
    nn = X.size // 3
    rp, ci = O.pattern(nn, conn)
    colors, nc = O.color(nn, conn)
    K, f = O.assemble(X, conn, mu, lam, colors, nc, u, rp, ci)
    return rp, ci, K, f


# ---------------------------------------------------------------------------- J clamp, inversion
def test_clamped_and_inverted_elements_match_oracle(fm, S, O):
    cfg = fm.configs.make(3, dims=(3, 3, 3))
    X, conn = cfg["X"], cfg["conn"]
    nn, ne = X.size // 3, conn.size // 4
    P = X.reshape(3, nn).T
    c = conn.reshape(4, ne).T
    u = np.random.default_rng(7).uniform(-0.01, 0.01, 3 * nn)
    # move an interior node towards the opposite face of one incident tet so that
    # that tet keeps 0.5% of its volume: det F = 0.005, inside the clamp
    v = 1 + 4 * 1 + 16 * 1  # grid node (1, 1, 1)
    e = int(np.where((c == v).any(axis=1))[0][0])
    others = [int(w) for w in c[e] if w != v]
    a, b, d = (P[w] + u.reshape(nn, 3)[w] for w in others)
    nrm = np.cross(b - a, d - a)
    nrm /= np.linalg.norm(nrm)
    xv = P[v] + u.reshape(nn, 3)[v]
    dist = np.dot(xv - a, nrm)
    target = xv - (1.0 - 0.005) * dist * nrm
    u.reshape(nn, 3)[v] = target - P[v]
    dF = det_F(X, conn, u)
    assert 0.004 < abs(dF[e]) < 0.006
    assert (np.abs(dF) < 0.01).sum() >= 1
    fm.configs.load(S, cfg)
    rp, ci, Ko, fo = oracle_assembly(O, X, conn, cfg["mu"], cfg["lam"], u)
    f64, K64 = S.assemble(u, fm.FM_FP64)
    assert rel(f64, fo) < 1e-12 and rel(K64, Ko) < 1e-12
    f32, K32 = S.assemble(u, fm.FM_MIXED)
    assert rel(f32, fo) < 1e-12 and rel(K32, Ko) < 2e-6
    # the element fields see the same clamp (energy with ln(max(J, 0.01)))
    fld = S.element_fields(u, fields=("det_F", "energy"))
    assert rel(fld["det_F"], dF) < 1e-12
    ofl = O.element_fields(X, conn, cfg["mu"], cfg["lam"], u)
    assert rel(fld["energy"], ofl["energy"]) < 1e-12


def test_inverted_element_region_matches_oracle(fm, S, O):
    # a node pushed through the opposite face: det F < 0 in that tet, and |det F|
    # above and below the clamp in its neighbours
    cfg = fm.configs.make(2, dims=(3, 3, 3))
    X, conn = cfg["X"], cfg["conn"]
    nn = X.size // 3
    u = np.zeros(3 * nn)
    v = 1 + 4 * 2 + 16 * 1
    u.reshape(nn, 3)[v] = (0.45 / 3, -0.4 / 3, 0.42 / 3)
    dF = det_F(X, conn, u)
    assert (dF < 0).sum() >= 1 and (np.abs(dF) < 0.01).sum() >= 1
    fm.configs.load(S, cfg)
    rp, ci, Ko, fo = oracle_assembly(O, X, conn, cfg["mu"], cfg["lam"], u)
    f64, K64 = S.assemble(u, fm.FM_FP64)
    assert rel(f64, fo) < 1e-12 and rel(K64, Ko) < 1e-12
    _, K32 = S.assemble(u, fm.FM_MIXED)
    assert rel(K32, Ko) < 2e-6


# ---------------------------------------------------------------------------- zero-volume skip
def test_zero_volume_element_contributes_nothing(fm, S, O):
    X, conn = fm.mesh_box(2, 2, 2)
    nn, ne = X.size // 3, conn.size // 4
    # one extra element on four coplanar nodes of the x = 0 face: volume 0
    g = lambda i, j, k: i + 3 * j + 9 * k  # noqa: E731
    flat = np.array([g(0, 0, 0), g(0, 1, 0), g(0, 0, 1), g(0, 1, 1)], np.uint32)
    c = conn.reshape(4, ne)
    conn2 = np.concatenate([c, flat[:, None]], axis=1).ravel()
    mu = np.full(ne + 1, 0.4)
    lam = np.full(ne + 1, 0.6)
    u = np.random.default_rng(3).uniform(-0.03, 0.03, 3 * nn)
    S.set_mesh(X, conn2)
    S.set_material(mu, lam)
    S.set_bcs(np.zeros(0, np.uint32))
    rp2, ci2, Ko2, fo2 = oracle_assembly(O, X, conn2, mu, lam, u)
    f64, K64 = S.assemble(u, fm.FM_FP64)
    assert rel(f64, fo2) < 1e-12 and rel(K64, Ko2) < 1e-12
    # against the mesh without it: the same force, the same tangent on the
    # common entries and zeros on the entries only the flat element adds
    rp, ci, Ko, fo = oracle_assembly(O, X, conn, mu[:ne], lam[:ne], u)
    assert rel(f64, fo) < 1e-12
    dense2 = {}
    for r in range(rp2.size - 1):
        for k in range(rp2[r], rp2[r + 1]):
            dense2[(r, int(ci2[k]))] = K64[k]
    common = set()
    for r in range(rp.size - 1):
        for k in range(rp[r], rp[r + 1]):
            common.add((r, int(ci[k])))
            assert abs(dense2[(r, int(ci[k]))] - Ko[k]) <= 1e-12 * np.abs(Ko).max()
    assert len(dense2) > len(common)
    assert all(dense2[key] == 0.0 for key in dense2 if key not in common)
    fld = S.element_fields(u, fields=("det_F", "von_mises", "energy"))
    assert fld["det_F"][ne] == 0.0 and fld["von_mises"][ne] == 0.0 and fld["energy"][ne] == 0.0


# ---------------------------------------------------------------------------- colour-launch assembly
def run_asm_child(env_asm, out):
    env = dict(os.environ)
    if env_asm:
        env["FM_ASM"] = env_asm
    else:
        env.pop("FM_ASM", None)
    r = subprocess.run([sys.executable, os.path.join(HERE, "asm_check.py"), str(out)], capture_output=True,
                       text=True, timeout=300, env=env)
    assert r.returncode == 0 and "ASM OK" in r.stdout, r.stdout + r.stderr
    return dict(np.load(out))


def test_color_assembly_matches_oracle_and_node_gather(fm, O, tmp_path):
    sys.path.insert(0, HERE)
    import asm_check

    col = run_asm_child("color", tmp_path / "color.npz")
    node = run_asm_child(None, tmp_path / "node.npz")
    cfg, u = asm_check.problem()
    rp, ci, Ko, fo = oracle_assembly(O, cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], u)
    for got in (col, node):
        assert rel(got["f64"], fo) < 1e-12 and rel(got["K64"], Ko) < 1e-12
        assert rel(got["f32"], fo) < 1e-12 and rel(got["K32"], Ko) < 2e-6
    # the node-gather kernel adds each entry's rounded element terms in colour
    # order, as the colour launches do: the same bits
    for k in ("f64", "K64", "f32", "K32", "R"):
        assert np.array_equal(col[k], node[k]), (k, rel(col[k], node[k]))
    # the load step (predictor lift, Dirichlet masking) against the oracle
    P = O.NewtonProblem(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], cfg["dir_dofs"], cfg["dir_vals"],
                        cfg["f_ext"], cfg["n_steps"], newton_rtol=1e-10, pcg_tol=1e-12, pcg_max=2000)
    uo = np.zeros(cfg["X"].size)
    rc, st = P.step(1, uo)
    for got in (col, node):
        assert got["newton1"][1] == 1 and got["newton1"][0] == st["newton_iters"]
        assert rel(got["u1"], uo) < 1e-9
        assert rel(got["u1_mixed"], uo) < 1e-7


def fan_mesh(m=40):
    """A ring of m points around a centre node with an apex above and below:
    2m tets share the centre node (more than the node-gather kernel's 32)."""
    ang = 2 * np.pi * np.arange(m) / m
    pts = [(0.0, 0.0, 0.0)] + [(np.cos(t), np.sin(t), 0.0) for t in ang] + [(0.0, 0.0, 0.7), (0.0, 0.0, -0.7)]
    top, bot = m + 1, m + 2
    tets = []
    for i in range(m):
        a, b = 1 + i, 1 + (i + 1) % m
        tets.append((0, a, b, top))
        tets.append((0, b, a, bot))
    P = np.array(pts)
    X = np.ascontiguousarray(P.T).ravel()
    conn = np.ascontiguousarray(np.array(tets, np.uint32).T).ravel()
    return X, conn


def test_color_assembly_fallback_for_high_valence_nodes(fm, S, O):
    X, conn = fan_mesh()
    nn, ne = X.size // 3, conn.size // 4
    assert np.bincount(conn, minlength=nn).max() > 32
    # the orientation gives positive volumes
    assert (det_F(X, conn, np.full(3 * nn, 0.0)) > 0).all()
    mu = np.random.default_rng(2).uniform(0.5, 1.5, ne)
    lam = 2 * mu
    u = np.random.default_rng(4).uniform(-0.02, 0.02, 3 * nn)
    S.set_mesh(X, conn)
    S.set_material(mu, lam)
    S.set_bcs(np.zeros(0, np.uint32))
    rp, ci, Ko, fo = oracle_assembly(O, X, conn, mu, lam, u)
    rpg, cig = S.pattern()
    assert np.array_equal(rpg, rp) and np.array_equal(cig, ci)
    f64, K64 = S.assemble(u, fm.FM_FP64)
    assert rel(f64, fo) < 1e-12 and rel(K64, Ko) < 1e-12
    _, K32 = S.assemble(u, fm.FM_MIXED)
    assert rel(K32, Ko) < 2e-6
    # a load step: ring clamped, apexes pulled apart along z
    ring = np.arange(1, nn - 2)
    dofs = np.sort(np.concatenate([3 * ring, 3 * ring + 1, 3 * ring + 2, [3 * (nn - 2) + 2, 3 * (nn - 1) + 2]]))
    vals = np.zeros(dofs.size)
    vals[dofs == 3 * (nn - 2) + 2] = 0.05
    vals[dofs == 3 * (nn - 1) + 2] = -0.05
    S.set_bcs(dofs.astype(np.uint32), vals)
    c = S.newton_cfg(n_steps=2, prec=fm.FM_FP64, newton_rtol=1e-10, pcg_tol=1e-12, pcg_max=500)
    P = O.NewtonProblem(X, conn, mu, lam, dofs, vals, None, 2, newton_rtol=1e-10, pcg_tol=1e-12, pcg_max=500)
    uo = np.zeros(3 * nn)
    for s in (1, 2):
        rc, st = S.newton_step(s, c)
        rco, sto = P.step(s, uo)
        assert st["converged"] and st["newton_iters"] == sto["newton_iters"]
    assert rel(S.get_u(), uo) < 1e-9


# ---------------------------------------------------------------------------- configuration workloads
def test_compression_steps_with_predictor_match_oracle(fm, S, O):
    # configuration 2 form at 8^3: jittered, lognormal mu, x=1 compressed 30% over 20 steps
    cfg = fm.configs.make(2, dims=(8, 8, 8))
    fm.configs.load(S, cfg)
    assert cfg["prec"] == fm.FM_FP64 and (cfg["dir_vals"] < 0).any()
    c = fm.configs.newton_cfg(cfg, pcg_tol=1e-12, pcg_max=3000)
    P = O.NewtonProblem(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], cfg["dir_dofs"], cfg["dir_vals"], None,
                        cfg["n_steps"], newton_rtol=cfg["newton_rtol"], pcg_tol=1e-12, pcg_max=3000)
    uo = np.zeros(cfg["X"].size)
    for s in (1, 2, 3):
        rc, st = S.newton_step(s, c)
        rco, sto = P.step(s, uo)
        assert st["converged"] and st["newton_iters"] == sto["newton_iters"]
        # the predictor solve ran (one factorization more than Newton iterations)
        assert st["factorizations"] == st["newton_iters"] + 1
    assert rel(S.get_u(), uo) < 1e-9
    # the mixed path on the same steps reaches the FP64 answer to its tolerance
    S.set_u(None)
    cm = fm.configs.newton_cfg(cfg, prec=fm.FM_MIXED, inner_tol=1e-4)
    for s in (1, 2, 3):
        rc, st = S.newton_step(s, cm)
        assert st["converged"]
    assert rel(S.get_u(), uo) < 1e-7


def test_beam_adjoint_traction_and_mu_match_oracle_and_fd(fm, S, O):
    # configuration 5 form at 16x2x2: gradient of the mean tip deflection with
    # respect to the traction magnitude and a direction of mu_e
    cfg = fm.configs.make(5, dims=(16, 2, 2))
    cfg["prec"] = fm.FM_FP64
    fm.configs.load(S, cfg)
    nn, ne = cfg["X"].size // 3, cfg["conn"].size // 4
    gi = np.arange(nn) % (cfg["dims"][0] + 1)
    tip = np.where(gi == cfg["dims"][0])[0]
    # each solve: Newton to 1e-10, then two full Newton iterations that land on
    # the FP64 floor of ||R|| (the central differences need the states to ~1e-14)
    c = fm.configs.newton_cfg(cfg, newton_rtol=1e-10, pcg_tol=1e-13, pcg_max=5000, max_newton=60)
    cpol = fm.configs.newton_cfg(cfg, newton_rtol=0.0, pcg_tol=1e-13, pcg_max=5000, max_newton=2)

    def solve(step):
        rc, st = S.newton_step(step, c)
        assert st["converged"]
        S.newton_step(step, cpol, check=False)

    solve(1)
    u1 = S.get_u()
    solve(2)
    u = S.get_u()
    lf = 2 / cfg["n_steps"]
    g = np.zeros(3 * nn)
    g.reshape(nn, 3)[tip, 2] = 1.0 / tip.size
    dmu, dlam, dload, its = S.adjoint(g, tol=1e-13, max_it=5000)
    P = O.NewtonProblem(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], cfg["dir_dofs"], cfg["dir_vals"],
                        cfg["f_ext"], cfg["n_steps"], newton_rtol=1e-13, pcg_tol=1e-13, pcg_max=5000, max_newton=60,
                        ls_mode=1)
    uo = np.zeros(3 * nn)
    for s in (1, 2):
        P.step(s, uo)
    assert rel(u, uo) < 1e-9
    dmo, dlo, dldo, _, _ = P.adjoint(uo, g, tol=1e-13, max_iter=5000)
    assert rel(dmu, dmo) < 1e-8 and rel(dlam, dlo) < 1e-8 and abs(dload - dldo) <= 1e-8 * abs(dldo)
    v = np.random.default_rng(3).standard_normal(ne) * cfg["mu"]

    def defl(scale=1.0, mu=None):
        S.set_load(scale * cfg["f_ext"])
        S.set_material(cfg["mu"] if mu is None else mu, cfg["lam"])
        S.set_u(u1)
        solve(2)
        return float(S.get_u().reshape(nn, 3)[tip, 2].mean())

    h = 1e-6
    fd_t = (defl(1 + h) - defl(1 - h)) / (2 * h)
    fd_mu = (defl(mu=cfg["mu"] + h * v) - defl(mu=cfg["mu"] - h * v)) / (2 * h)
    ad_t, ad_mu = dload * lf, float(dmu @ v)
    assert abs(ad_t - fd_t) <= 1e-5 * abs(fd_t), (ad_t, fd_t)
    assert abs(ad_mu - fd_mu) <= 1e-5 * abs(fd_mu), (ad_mu, fd_mu)
