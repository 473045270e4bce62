"""Pins the FP64 CPU oracle: golden vectors from the compiled reference,
closed forms, finite differences and brute force on tiny inputs.

Each check is chosen so that a plausible slip in the oracle (a dropped term, a
transposed operand, a wrong index) fails it; none of them retypes the oracle's
own formula.
"""
import itertools

import numpy as np
import pytest

from conftest import golden


def lame(E, nu):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


# ----------------------------------------------------------------------------- a5 stress
def test_pk1_matches_reference_golden(O):
    g = golden("pk1.npz")
    for F, E, nu, P in zip(g["F"], g["E"], g["nu"], g["P"]):
        mu, lam = lame(E, nu)
        scale = max(np.abs(P).max(), mu)
        # literal restatement: same operations as compute_pk1_stress
        assert np.abs(O.pk1_refform(F, mu, lam) - P).max() <= 4e-15 * scale
        # closed form used by the kernels: P = mu F + (lam lnJ - mu) F^-T
        assert np.abs(O.pk1(F, mu, lam) - P).max() <= 1e-13 * scale


def test_pk1_survey_probe_value(O):
    F = np.array([[1.1, 0.05, 0.0], [-0.02, 0.95, 0.03], [0.01, 0.0, 1.05]])
    mu, lam = lame(1.0, 0.3)
    assert abs(O.pk1(F, mu, lam)[0] - 1.229017653728824e-01) < 1e-15


def test_pk1_is_energy_derivative_and_objective(O):
    rng = np.random.default_rng(1)
    mu, lam = 0.7, 2.1
    for _ in range(10):
        F = np.eye(3) + rng.uniform(-0.25, 0.25, (3, 3))
        P = O.pk1(F, mu, lam).reshape(3, 3)
        h = 1e-6
        Pfd = np.zeros((3, 3))
        for i, j in itertools.product(range(3), range(3)):
            Fp, Fm = F.copy(), F.copy()
            Fp[i, j] += h
            Fm[i, j] -= h
            Pfd[i, j] = (O.energy(Fp, mu, lam) - O.energy(Fm, mu, lam)) / (2 * h)
        assert np.abs(P - Pfd).max() < 1e-8
        # objectivity P(QF) = Q P(F)
        A = rng.standard_normal((3, 3))
        Q, _ = np.linalg.qr(A)
        if np.linalg.det(Q) < 0:
            Q[:, 0] *= -1
        assert np.abs(O.pk1(Q @ F, mu, lam).reshape(3, 3) - Q @ P).max() < 1e-13
    assert np.abs(O.pk1(np.eye(3), mu, lam)).max() == 0.0
    # small-strain limit: P -> lam tr(eps) I + 2 mu eps
    eps = 1e-7 * rng.standard_normal((3, 3))
    eps = 0.5 * (eps + eps.T)
    P = O.pk1(np.eye(3) + eps, mu, lam).reshape(3, 3)
    lin = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    assert np.abs(P - lin).max() < 1e-12


# ----------------------------------------------------------------------------- a2 mesh
@pytest.mark.parametrize("N", [1, 2, 4, 8])
def test_mesher_numbering_matches_reference(O, N):
    g = golden("mesh_pattern_color.npz")
    X, conn = O.box_mesh(N, N, N)
    assert np.array_equal(X, g[f"ref_X_{N}"])  # same node ids and coordinates
    assert conn.size == g[f"ref_conn_{N}"].size  # same tet count


def _tet_volumes(X, conn):
    nn, ne = X.size // 3, conn.size // 4
    P = X.reshape(3, nn).T[conn.reshape(4, ne).T]  # ne x 4 x 3
    J = np.stack([P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]], axis=2)
    return np.linalg.det(J) / 6.0


@pytest.mark.parametrize("dims", [(3, 3, 3), (4, 2, 3)])
def test_mesher_conforming_positive(O, dims):
    nx, ny, nz = dims
    X, conn = O.box_mesh(nx, ny, nz, 2.0, 1.0, 1.5)
    ne = conn.size // 4
    v = _tet_volumes(X, conn)
    assert v.min() > 0
    assert abs(v.sum() - 3.0) < 1e-12
    faces = {}
    c = conn.reshape(4, ne).T
    for e in range(ne):
        for skip in range(4):
            f = tuple(sorted(np.delete(c[e], skip)))
            faces[f] = faces.get(f, 0) + 1
    counts = np.bincount(list(faces.values()))
    n_bnd = 2 * 2 * (nx * ny + ny * nz + nx * nz)
    assert counts[1] == n_bnd and counts.size == 3


# ----------------------------------------------------------------------------- a10 pattern
def _edges_kuhn(N):
    return 3 * N * (N + 1) ** 2 + 3 * N * N * (N + 1) + N ** 3


@pytest.mark.parametrize("N", [1, 2, 4, 8])
def test_pattern_rows_match_reference_and_closed_form(O, N):
    g = golden("mesh_pattern_color.npz")
    X, conn = O.box_mesh(N, N, N)
    nn = X.size // 3
    rp, ci = O.pattern(nn, conn)
    assert np.array_equal(rp, g[f"ref_rowptr_kuhn_{N}"])
    assert rp[-1] == 9 * (nn + 2 * _edges_kuhn(N))
    # the reference's own (non-conforming) mesh through the same rule
    Xr, cr = g[f"ref_X_{N}"], g[f"ref_conn_{N}"]
    rpr, _ = O.pattern(nn, cr)
    assert np.array_equal(rpr, g[f"ref_rowptr_refmesh_{N}"])
    for r in range(rp.size - 1):
        row = ci[rp[r]:rp[r + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0)


def test_pattern_brute_force_pairs(O):
    X, conn = O.box_mesh(2, 3, 2)
    nn, ne = X.size // 3, conn.size // 4
    rp, ci = O.pattern(nn, conn)
    want = set()
    for e in range(ne):
        nodes = conn.reshape(4, ne)[:, e]
        for a, b in itertools.product(nodes, nodes):
            for i, k in itertools.product(range(3), range(3)):
                want.add((3 * int(a) + i, 3 * int(b) + k))
    got = {(r, int(c)) for r in range(3 * nn) for c in ci[rp[r]:rp[r + 1]]}
    assert got == want


# ----------------------------------------------------------------------------- a13 coloring
@pytest.mark.parametrize("N", [1, 2, 4, 8])
def test_coloring_matches_reference(O, N):
    g = golden("mesh_pattern_color.npz")
    X, conn = O.box_mesh(N, N, N)
    colors, nc = O.color(X.size // 3, conn)
    assert nc == int(g[f"ref_ncolors_kuhn_{N}"][0])
    assert np.array_equal(colors, g[f"ref_colors_kuhn_{N}"])
    if f"ref_colors_refmesh_{N}" in g:
        cr = g[f"ref_conn_{N}"]
        colr, ncr = O.color(X.size // 3, cr)
        assert ncr == int(g[f"ref_ncolors_refmesh_{N}"][0])
        assert np.array_equal(colr, g[f"ref_colors_refmesh_{N}"])


def test_coloring_first_fit_invariant(O):
    X, conn = O.box_mesh(3, 2, 4)
    nn, ne = X.size // 3, conn.size // 4
    colors, nc = O.color(nn, conn)
    c = conn.reshape(4, ne).T
    for e in range(ne):
        earlier = [f for f in range(e) if set(c[f]) & set(c[e])]
        used = {int(colors[f]) for f in earlier}
        assert int(colors[e]) not in used
        assert all(k in used for k in range(int(colors[e])))


# ----------------------------------------------------------------------------- a9 element
def _rand_tet(rng):
    Xe = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1]], float) * 0.4 + rng.uniform(-0.04, 0.04, (4, 3))
    return Xe


def test_element_tangent_is_force_derivative(O):
    rng = np.random.default_rng(2)
    for _ in range(6):
        Xe = _rand_tet(rng)
        ue = rng.uniform(-0.06, 0.06, 12)
        mu, lam = np.exp(0.3 * rng.standard_normal()), 49 * np.exp(0.3 * rng.standard_normal())
        f, K, skipped = O.element(Xe, ue, mu, lam)
        assert skipped == 0
        h = 1e-6
        Kfd = np.zeros((12, 12))
        for j in range(12):
            up, um = ue.copy(), ue.copy()
            up[j] += h
            um[j] -= h
            Kfd[:, j] = (O.element(Xe, up, mu, lam)[0] - O.element(Xe, um, mu, lam)[0]) / (2 * h)
        assert np.abs(K - Kfd).max() < 1e-7 * np.abs(K).max()
        assert np.abs(K - K.T).max() < 1e-14 * np.abs(K).max()
        # rigid translation: no force change
        t = np.tile(rng.standard_normal(3), 4)
        assert np.abs(K @ t).max() < 1e-12 * np.abs(K).max()
        assert np.abs(f.reshape(4, 3).sum(axis=0)).max() < 1e-14 * max(1, np.abs(f).max())


def test_element_tangent_at_the_j_clamp_is_force_derivative(O):
    """J < 0.01 (compressed and inverted): ln J is clamped at ln 0.01
    (FEM/constitutive_models.hpp:96), so the force has no lambda ln J
    derivative and the tangent drops the lam' term.  Central FD of f pins it:
    a tangent that kept lam' = lam would differ by lam V (F^-T g)(F^-T g)^T,
    ~1e4 times the tolerance at det F = 0.005."""
    rng = np.random.default_rng(12)
    for target in (0.005, -0.004, 0.0095):
        Xe = _rand_tet(rng)
        # move node 3 toward the plane of nodes 0..2: det F = target
        n = np.cross(Xe[1] - Xe[0], Xe[2] - Xe[0])
        n /= np.linalg.norm(n)
        d = np.dot(Xe[3] - Xe[0], n)
        ue = np.zeros(12)
        ue[9:12] = -(1.0 - target) * d * n
        ue += rng.uniform(-1e-4, 1e-4, 12) * (np.arange(12) < 9)
        J0 = np.stack([Xe[1] - Xe[0], Xe[2] - Xe[0], Xe[3] - Xe[0]], axis=1)
        x = Xe + ue.reshape(4, 3)
        J1 = np.stack([x[1] - x[0], x[2] - x[0], x[3] - x[0]], axis=1)
        detF = np.linalg.det(J1) / np.linalg.det(J0)
        assert abs(detF) < 0.01 and abs(abs(detF) - abs(target)) < 1e-3
        mu, lam = 0.7, 34.0
        f, K, skipped = O.element(Xe, ue, mu, lam)
        assert skipped == 0
        h = 1e-7
        Kfd = np.zeros((12, 12))
        for j in range(12):
            up, um = ue.copy(), ue.copy()
            up[j] += h
            um[j] -= h
            Kfd[:, j] = (O.element(Xe, up, mu, lam)[0] - O.element(Xe, um, mu, lam)[0]) / (2 * h)
        assert np.abs(K - Kfd).max() < 1e-6 * np.abs(K).max(), (target, np.abs(K - Kfd).max() / np.abs(K).max())
        assert np.abs(K - K.T).max() < 1e-13 * np.abs(K).max()


def test_element_linear_limit_is_BtDB(O):
    rng = np.random.default_rng(3)
    Xe = _rand_tet(rng)
    mu, lam = 0.8, 1.7
    f, K, _ = O.element(Xe, np.zeros(12), mu, lam)
    assert np.abs(f).max() < 1e-15 * np.abs(K).max()  # FMA rounding of mu g - mu g
    J = np.stack([Xe[1] - Xe[0], Xe[2] - Xe[0], Xe[3] - Xe[0]], axis=1)
    V = abs(np.linalg.det(J)) / 6
    G = np.linalg.solve(J.T, np.array([[-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]], float)).T  # 4x3
    B = np.zeros((6, 12))
    for a in range(4):
        gx, gy, gz = G[a]
        B[:, 3 * a:3 * a + 3] = [[gx, 0, 0], [0, gy, 0], [0, 0, gz], [gy, gx, 0], [0, gz, gy], [gz, 0, gx]]
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[3:, 3:] = mu * np.eye(3)
    assert np.abs(K - V * B.T @ D @ B).max() < 1e-13 * np.abs(K).max()
    w = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(w) < 1e-12 * w.max()) == 6


def test_element_degenerate_is_skipped(O):
    Xe = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float)
    f, K, skipped = O.element(Xe, np.ones(12), 1.0, 1.0)
    assert skipped == 1 and not f.any() and not K.any()


# ----------------------------------------------------------------------------- a11 assembly
def test_assembly_matches_dense_brute_force(O):
    rng = np.random.default_rng(4)
    X, conn = O.box_mesh(2, 1, 2)
    nn, ne = X.size // 3, conn.size // 4
    X = X + rng.uniform(-0.05, 0.05, X.size)
    rp, ci = O.pattern(nn, conn)
    colors, nc = O.color(nn, conn)
    mu = np.exp(0.3 * rng.standard_normal(ne))
    lam = 1.5 * mu
    u = rng.uniform(-0.03, 0.03, 3 * nn)
    K, f = O.assemble(X, conn, mu, lam, colors, nc, u, rp, ci)
    Kd = np.zeros((3 * nn, 3 * nn))
    fd = np.zeros(3 * nn)
    c = conn.reshape(4, ne).T
    for e in range(ne):
        Xe = X.reshape(3, nn).T[c[e]]
        ue = u.reshape(nn, 3)[c[e]].ravel()
        fe, Ke, _ = O.element(Xe, ue, mu[e], lam[e])
        dofs = (3 * c[e][:, None] + np.arange(3)).ravel()
        Kd[np.ix_(dofs, dofs)] += Ke
        fd[dofs] += fe
    Ks = np.zeros_like(Kd)
    for r in range(3 * nn):
        Ks[r, ci[rp[r]:rp[r + 1]]] = K[rp[r]:rp[r + 1]]
    assert np.abs(Ks - Kd).max() < 1e-14 * np.abs(Kd).max()
    assert np.abs(f - fd).max() < 1e-15 * np.abs(fd).max() * 10
    assert np.abs(Ks - Ks.T).max() < 1e-14 * np.abs(Kd).max()
    # threaded assembly gives the same bits (colour order is fixed)
    O.set_threads(4)
    K4, f4 = O.assemble(X, conn, mu, lam, colors, nc, u, rp, ci)
    O.set_threads(1)
    assert np.array_equal(K4, K) and np.array_equal(f4, f)


# ----------------------------------------------------------------------------- a14 ILU(0)
def _sort_rows(rp, ci, va):
    perm = np.concatenate([rp[r] + np.argsort(ci[rp[r]:rp[r + 1]], kind="stable") for r in range(rp.size - 1)])
    return ci[perm], va[perm], perm


@pytest.mark.parametrize("name", ["poisson16_refcols", "poisson16_sorted", "elastic2"])
def test_ilu0_matches_reference_golden(O, name):
    g = golden("ilu0.npz")
    rp, ci, va = g[f"{name}_rp"], g[f"{name}_ci"], g[f"{name}_val"]
    cs, vs, perm = _sort_rows(rp, ci, va)
    LU, diag = O.ilu0(rp, cs, vs)
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    Lref, Uref = g[f"{name}_L"][perm], g[f"{name}_U"][perm]
    low, up = cs < rows, cs > rows
    scale = np.abs(va).max()
    assert np.abs(LU[low] - Lref[low]).max() <= 1e-15 * max(1.0, np.abs(Lref[low]).max())
    assert np.abs(LU[up] - Uref[up]).max() <= 2e-16 * scale
    assert np.abs(diag - g[f"{name}_diag"]).max() <= 2e-16 * scale
    y = O.ilu0_apply(rp, cs, LU, diag, g[f"{name}_x"])
    assert np.abs(y - g[f"{name}_y"]).max() <= 1e-14 * np.abs(g[f"{name}_y"]).max()


def test_ilu0_defining_identity(O):
    """(L U)_ij = a_ij on the pattern of A, with L unit lower."""
    g = golden("ilu0.npz")
    rp, ci, va = g["elastic2_rp"], g["elastic2_ci"], g["elastic2_val"]
    LU, diag = O.ilu0(rp, ci, va)
    n = rp.size - 1
    Ld, Ud, Ad = np.eye(n), np.zeros((n, n)), np.zeros((n, n))
    for r in range(n):
        for k in range(rp[r], rp[r + 1]):
            c = ci[k]
            Ad[r, c] = va[k]
            if c < r:
                Ld[r, c] = LU[k]
            else:
                Ud[r, c] = LU[k]
    M = Ld @ Ud
    mask = Ad != 0
    assert np.abs(M[mask] - Ad[mask]).max() < 1e-13 * np.abs(Ad).max()


def test_ilu0_pivot_perturbation(O):
    rp = np.array([0, 2, 4], np.uint32)
    ci = np.array([0, 1, 0, 1], np.uint32)
    va = np.array([1.0, 1.0, 1.0, 1.0])  # u_11 = 1 - 1 = 0 -> +1e-8
    LU, diag = O.ilu0(rp, ci, va)
    assert diag[1] == 1e-8 and diag[0] == 1.0


# ----------------------------------------------------------------------------- a16 PCG
@pytest.mark.parametrize("gn", [16, 32, 48, 96])
def test_pcg_matches_reference_golden(O, gn):
    """Manufactured 2-D Poisson of TESTS/external_validation_runner.cpp:65-157
    (tol 1e-10, max 4000): the compiled reference takes 16/27/37/64 iterations
    for 16^2/32^2/48^2/96^2 unknowns, with either column order."""
    import make_golden as MG
    g = golden("pcg.npz")
    assert int(g[f"poisson{gn}_refcols_iters"][0]) == int(g[f"poisson{gn}_sorted_iters"][0])
    rp, ci, va, b, xs = MG.poisson(gn, True)
    x, its, rel = O.pcg(rp, ci, va, b, tol=1e-10, max_iter=4000)
    assert its == int(g[f"poisson{gn}_sorted_iters"][0])
    xr = g[f"poisson{gn}_sorted_x"]
    assert np.abs(x - xr).max() <= 1e-13 * np.abs(xr).max()
    assert np.abs(x - xs).max() < 1e-6


def test_pcg_elastic_golden_and_dense_solve(O):
    g = golden("pcg.npz")
    rp, ci, va, b = g["elastic2_rp"], g["elastic2_ci"], g["elastic2_val"], g["elastic2_b"]
    x, its, rel = O.pcg(rp, ci, va, b, tol=1e-10, max_iter=300)
    assert its == int(g["elastic2_iters"][0])
    assert np.abs(x - g["elastic2_x"]).max() <= 1e-12 * np.abs(x).max()
    n = rp.size - 1
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[rp[r]:rp[r + 1]]] = va[rp[r]:rp[r + 1]]
    xd = np.linalg.solve(A, b)
    assert np.abs(x - xd).max() < 1e-8 * np.abs(xd).max()


# ----------------------------------------------------------------------------- a18 Newton
def _uniaxial_lateral_stretch(l1, mu, lam):
    """lateral stretch of homogeneous uniaxial tension with free lateral faces:
    mu (l2 - 1/l2) + lam ln(l1 l2^2) / l2 = 0"""
    lo, hi = 0.3, 1.5
    f = lambda l2: mu * (l2 - 1 / l2) + lam * np.log(l1 * l2 * l2) / l2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(lo) * f(mid) <= 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def test_newton_homogeneous_uniaxial_closed_form(O):
    N = 3
    X, conn = O.box_mesh(N, N, N)
    nn, ne = X.size // 3, conn.size // 4
    mu, lam = lame(1.0, 0.3)
    x, y, z = X[:nn], X[nn:2 * nn], X[2 * nn:]
    dofs, vals = [], []
    delta = 0.2
    for n in range(nn):
        if x[n] == 0.0:
            dofs.append(3 * n); vals.append(0.0)
        if x[n] == 1.0:
            dofs.append(3 * n); vals.append(delta)
        if y[n] == 0.0:
            dofs.append(3 * n + 1); vals.append(0.0)
        if z[n] == 0.0:
            dofs.append(3 * n + 2); vals.append(0.0)
    P = O.NewtonProblem(X, conn, np.full(ne, mu), np.full(ne, lam), dofs, vals, None, 4, newton_rtol=1e-12,
                        pcg_tol=1e-12, pcg_max=500)
    u = np.zeros(3 * nn)
    for s in range(1, 5):
        rc, st = P.step(s, u)
        assert rc == 0 and st["converged"]
    l2 = _uniaxial_lateral_stretch(1 + delta, mu, lam)
    assert np.abs(u[0::3] - delta * x).max() < 1e-10
    assert np.abs(u[1::3] - (l2 - 1) * y).max() < 1e-10
    assert np.abs(u[2::3] - (l2 - 1) * z).max() < 1e-10


# ----------------------------------------------------------------------------- a20 Lie
def test_rotation_log_exp_closed_forms(O):
    g = golden("lie.npz")
    for w, Eref, Lref in zip(g["w"], g["expm"], g["logm"]):
        R = O.rot_exp(w)
        th = np.linalg.norm(w)
        n = w / th
        Kn = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
        Rod = np.eye(3) + np.sin(th) * Kn + (1 - np.cos(th)) * Kn @ Kn
        assert np.abs(R - Rod).max() < 2e-16 * 8
        assert np.abs(R.T @ R - np.eye(3)).max() < 1e-15
        assert np.abs(O.rot_log(R) - w).max() < 1e-15 * max(1, 8 * th)
        # the reference Pade[3/3] exp (CORE/lie_operator.cpp:94-152) agrees only to
        # its own accuracy (~1e-9 at 0.5 rad, ~1e-7 at pi/2)
        assert np.abs(Eref.reshape(3, 3) - R).max() < 1e-6 * max(1.0, th)
    for th in (1e-12, 1e-6, 0.3, 2.0, np.pi - 1e-3, np.pi - 1e-7):
        n = np.array([0.3, -0.5, 0.81])
        n /= np.linalg.norm(n)
        w = th * n
        assert np.abs(O.rot_log(O.rot_exp(w)) - w).max() < 1e-13 * max(th, 1e-12) + 1e-15


def test_polar_rotation(O):
    rng = np.random.default_rng(6)
    for _ in range(20):
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        if np.linalg.det(Q) < 0:
            Q[:, 0] *= -1
        A = rng.standard_normal((3, 3))
        Ssym = np.eye(3) + 0.3 * (A @ A.T)
        R, st = O.polar(Q @ Ssym)
        assert st == 0 and np.abs(R - Q).max() < 1e-13


# ----------------------------------------------------------------------------- a22 adjoint
def test_adjoint_matches_central_fd(O):
    rng = np.random.default_rng(8)
    N = 2
    X, conn = O.box_mesh(N, N, N)
    nn, ne = X.size // 3, conn.size // 4
    x, z = X[:nn], X[2 * nn:]
    dofs, vals = [], []
    for n in range(nn):
        if x[n] == 0.0:
            dofs += [3 * n, 3 * n + 1, 3 * n + 2]; vals += [0.0, 0.0, 0.0]
        if x[n] == 1.0:
            dofs.append(3 * n); vals.append(0.1)
    f_ext = np.zeros(3 * nn)
    f_ext[2::3][x == 1.0] = -0.002
    mu0 = np.exp(0.3 * rng.standard_normal(ne))
    lam0 = 1.5 * np.exp(0.3 * rng.standard_normal(ne))
    top = np.where(z == 1.0)[0]
    ustar = rng.uniform(-0.01, 0.01, 3 * nn)

    def solve(mu, lam, scale=1.0):
        P = O.NewtonProblem(X, conn, mu, lam, dofs, scale * np.array(vals), scale * f_ext, 2, newton_rtol=1e-13,
                            pcg_tol=1e-13, pcg_max=1000)
        u = np.zeros(3 * nn)
        for s in (1, 2):
            rc, _ = P.step(s, u)
            assert rc == 0
        return P, u

    def loss(u):
        d = (u - ustar).reshape(nn, 3)[top]
        return 0.5 * np.sum(d * d)

    P, u = solve(mu0, lam0)
    dLdu = np.zeros(3 * nn)
    dLdu.reshape(nn, 3)[top] = (u - ustar).reshape(nn, 3)[top]
    dmu, dlam, dload, _, _ = P.adjoint(u, dLdu)
    h = 1e-6
    for _ in range(3):
        v, w = rng.standard_normal(ne), rng.standard_normal(ne)
        fd = (loss(solve(mu0 + h * v, lam0 + h * w)[1]) - loss(solve(mu0 - h * v, lam0 - h * w)[1])) / (2 * h)
        ad = dmu @ v + dlam @ w
        assert abs(fd - ad) <= 1e-5 * abs(ad) + 1e-12
    fd = (loss(solve(mu0, lam0, 1 + h)[1]) - loss(solve(mu0, lam0, 1 - h)[1])) / (2 * h)
    assert abs(fd - dload) <= 1e-5 * abs(dload) + 1e-12


def _slender_beam(O, nx=16, load=0.0036):
    # 8 x 1 x 1 cantilever, clamped at x = 0, transverse nodal load on the x = 8 face
    X, conn = O.box_mesh(nx, 2, 2, 8.0, 1.0, 1.0)
    nn, ne = X.size // 3, conn.size // 4
    x = X[:nn]
    clamp = np.where(x == 0.0)[0]
    dofs = np.sort(np.concatenate([3 * clamp, 3 * clamp + 1, 3 * clamp + 2])).astype(np.uint32)
    f_ext = np.zeros(3 * nn)
    tip = np.where(x == 8.0)[0]
    f_ext[3 * tip + 2] = -load / tip.size
    E, nu = 1.0, 0.45
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    return X, conn, np.full(ne, mu), np.full(ne, lam), dofs, np.zeros(dofs.size), f_ext


def test_total_energy_gradient_is_minus_residual(O):
    X, conn, mu, lam, dofs, vals, f_ext = _slender_beam(O, nx=6)
    nn = X.size // 3
    P = O.NewtonProblem(X, conn, mu, lam, dofs, vals, f_ext, 1)
    assert P.energy(1.0, np.zeros(3 * nn)) == 0.0  # psi(I) = 0 and no work at u = 0
    rng = np.random.default_rng(7)
    u = 0.05 * rng.standard_normal(3 * nn)
    u[dofs] = 0.0
    rp, ci = O.pattern(nn, conn)
    col, nc = O.color(nn, conn)
    _, f = O.assemble(X, conn, mu, lam, col, nc, u, rp, ci, want_K=False)
    lf = 0.7
    R = lf * f_ext - f
    R[dofs] = 0.0
    v = rng.standard_normal(3 * nn)
    v[dofs] = 0.0
    h = 1e-5
    fd = (P.energy(lf, u + h * v) - P.energy(lf, u - h * v)) / (2 * h)
    assert abs(fd + R @ v) <= 1e-7 * np.abs(R).sum() * np.abs(v).max()
    # a plausible slip (forgetting the work term) fails the same check
    Pn = O.NewtonProblem(X, conn, mu, lam, dofs, vals, None, 1)
    fdn = (Pn.energy(lf, u + h * v) - Pn.energy(lf, u - h * v)) / (2 * h)
    assert abs(fdn + R @ v) > 1e3 * abs(fd + R @ v)


def test_energy_line_search_converges_where_residual_rule_stalls(O):
    # Slender cantilever under a transverse load: the full Newton step raises
    # ||R|| by ~100x before quadratic convergence, so the residual-norm Armijo
    # rule backtracks into many short steps; the energy rule accepts it.
    X, conn, mu, lam, dofs, vals, f_ext = _slender_beam(O)
    nn = X.size // 3
    out = {}
    for mode in (0, 1):
        P = O.NewtonProblem(X, conn, mu, lam, dofs, vals, f_ext, 40, newton_rtol=1e-8, pcg_tol=1e-10,
                            pcg_max=3000, max_newton=60, ls_mode=mode)
        u = np.zeros(3 * nn)
        its = []
        for s in (1, 2):
            rc, st = P.step(s, u)
            assert rc == 0
            its.append(st["newton_iters"])
        out[mode] = (u, its, P.energy(2 / 40, u))
    assert max(out[1][1]) <= 5 and min(out[0][1]) > max(out[1][1])
    u0, u1 = out[0][0], out[1][0]
    assert np.abs(u1 - u0).max() <= 1e-6 * np.abs(u0).max()
    assert abs(out[1][2] - out[0][2]) <= 1e-9 * abs(out[0][2])


# ----------------------------------------------------------------------------- element fields (next row 3)
def test_sym_eig3_matches_lapack(O):
    rng = np.random.default_rng(7)
    mats = [np.eye(3) * 2.0, np.diag([1.0, 1.0, 3.0]), np.diag([3.0, 3.0, 1.0]), np.zeros((3, 3)),
            np.outer([1.0, 2.0, -1.0], [1.0, 2.0, -1.0])]
    for _ in range(40):
        A = rng.standard_normal((3, 3)) * 10.0 ** rng.uniform(-3, 3)
        mats.append(A + A.T)
    for _ in range(10):  # nearly repeated eigenvalues
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        mats.append(Q @ np.diag([1.0, 1.0 + 1e-9, -2.0]) @ Q.T)
    for A in mats:
        w = O.sym_eig3(A)
        ref = np.sort(np.linalg.eigvalsh(A))[::-1]
        assert np.abs(w - ref).max() <= 1e-14 * max(np.abs(ref).max(), 1e-300) + 1e-300
    # only the upper triangle is read
    A = np.array([[1.0, 2.0, 3.0], [99.0, 4.0, 5.0], [99.0, 99.0, 6.0]])
    S = np.triu(A) + np.triu(A, 1).T
    assert np.abs(O.sym_eig3(A) - np.sort(np.linalg.eigvalsh(S))[::-1]).max() < 1e-13


def test_cauchy_and_von_mises_match_reference_golden(O):
    g = golden("stress.npz")
    for F, E, nu, S, vm in zip(g["F"], g["E"], g["nu"], g["sigma"], g["von_mises"]):
        mu, lam = lame(E, nu)
        scale = max(np.abs(S).max(), mu)
        assert np.abs(O.cauchy(F, mu, lam).ravel() - S).max() <= 1e-13 * scale
        assert abs(O.von_mises(S) - vm) <= 1e-14 * max(vm, scale)


def test_cauchy_closed_forms(O):
    mu, lam = 0.7, 2.1
    # uniaxial-type diagonal F: sigma_ii = (mu (a_i^2 - 1) + lam ln J) / J, off-diagonals zero
    for a in ([1.2, 0.9, 0.9], [0.8, 1.1, 1.05], [1.0, 1.0, 1.0]):
        F = np.diag(a)
        J = np.prod(a)
        s = O.cauchy(F, mu, lam)
        ref = np.diag([(mu * (ai * ai - 1.0) + lam * np.log(J)) / J for ai in a])
        assert np.abs(s - ref).max() < 1e-15
    rng = np.random.default_rng(3)
    for _ in range(10):
        F = np.eye(3) + rng.uniform(-0.25, 0.25, (3, 3))
        J = np.linalg.det(F)
        # sigma = (mu (F F^T - I) + lam ln J I) / J, symmetric
        ref = (mu * (F @ F.T - np.eye(3)) + lam * np.log(J) * np.eye(3)) / J
        s = O.cauchy(F, mu, lam)
        assert np.abs(s - ref).max() < 1e-14
        # objectivity: sigma(QF) = Q sigma(F) Q^T
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        if np.linalg.det(Q) < 0:
            Q[:, 0] *= -1
        assert np.abs(O.cauchy(Q @ F, mu, lam) - Q @ s @ Q.T).max() < 1e-14
        # von Mises is invariant too, and zero for a pure pressure
        assert abs(O.von_mises(Q @ s @ Q.T) - O.von_mises(s)) < 1e-13
    assert np.abs(O.cauchy(Q, mu, lam)).max() < 1e-14 * mu  # rotation: rounding only
    assert O.von_mises(np.eye(3) * 5.0) == 0.0
    # pure shear tau: sqrt(3) tau; uniaxial s: |s|
    tau = 0.3
    assert abs(O.von_mises(np.array([[0, tau, 0], [tau, 0, 0], [0, 0, 0]])) - np.sqrt(3) * tau) < 1e-15
    assert abs(O.von_mises(np.diag([-2.0, 0.0, 0.0])) - 2.0) < 1e-15


def test_element_fields_homogeneous_deformation(O):
    """u = G X gives F = I + G in every element: each field has a closed form."""
    import flexmesh as fm

    cfg = fm.configs.make(2, dims=(3, 2, 2))  # jittered interior, lognormal mu
    X = cfg["X"].reshape(3, -1)
    nn, ne = X.shape[1], cfg["conn"].size // 4
    G = np.array([[0.1, 0.02, -0.03], [0.01, -0.05, 0.04], [0.0, 0.03, 0.08]])
    u = (G @ X).T.ravel()
    f = O.element_fields(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], u)
    F = np.eye(3) + G
    J = np.linalg.det(F)
    C = F.T @ F
    assert np.abs(f["F"] - F.ravel()[:, None]).max() < 1e-14
    assert np.abs(f["det_F"] - J).max() < 1e-14
    assert np.abs(f["max_stretch"] - np.sqrt(max(np.linalg.eigvalsh(C).max(), 1.0))).max() < 1e-14
    for e in range(0, ne, 7):
        mu, lam = cfg["mu"][e], cfg["lam"][e]
        s = (mu * (F @ F.T - np.eye(3)) + lam * np.log(J) * np.eye(3)) / J
        c6 = f["cauchy"][:, e]
        assert np.abs(c6 - [s[0, 0], s[1, 1], s[2, 2], s[1, 2], s[0, 2], s[0, 1]]).max() < 1e-14
        assert abs(f["pressure"][e] - np.trace(s) / 3) < 1e-14
        assert abs(f["max_principal"][e] - np.linalg.eigvalsh(s).max()) < 1e-14
        dev = s - np.trace(s) / 3 * np.eye(3)
        assert abs(f["von_mises"][e] - np.sqrt(1.5 * np.sum(dev * dev))) < 1e-14
        lnJ = np.log(J)
        assert abs(f["energy"][e] - (0.5 * mu * (np.sum(F * F) - 3) - mu * lnJ + 0.5 * lam * lnJ ** 2)) < 1e-14
    # zero displacement: F = I, every stress-like field zero, stretch one
    z = O.element_fields(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], None)
    assert np.abs(z["cauchy"]).max() == 0.0 and np.abs(z["energy"]).max() == 0.0
    assert np.all(z["max_stretch"] == 1.0) and np.all(z["det_F"] == 1.0)
    assert nn == cfg["X"].size // 3


# ----------------------------------------------------------------------------- ZZ estimator (next row 4)
def zz_numpy(X, conn, u):
    """Independent restatement with numpy: F per element from the P1 gradients,
    volume-weighted node averages, least-squares patch fits (lstsq) in node-relative
    coordinates, constant fit where the linear fit is rank-deficient."""
    X = X.reshape(3, -1)
    nn, ne = X.shape[1], conn.size // 4
    cn = conn.reshape(4, ne)
    uu = u.reshape(-1, 3)
    g = np.zeros((ne, 3))
    vol = np.zeros(ne)
    cen = np.zeros((ne, 3))
    for e in range(ne):
        P = X[:, cn[:, e]].T
        D = (P[1:] - P[0]).T
        vol[e] = abs(np.linalg.det(D)) / 6
        cen[e] = P.mean(axis=0)
        U = uu[cn[:, e]]
        Gu = (U[1:] - U[0]).T @ np.linalg.inv(D)
        g[e] = np.diag(Gu)
    raw = np.zeros((nn, 3))
    rec = np.zeros((nn, 3))
    for v in range(nn):
        es = np.where((cn == v).any(axis=0))[0]
        raw[v] = (g[es] * vol[es, None]).sum(0) / vol[es].sum()
        A = np.hstack([np.ones((es.size, 1)), cen[es] - X[:, v]])
        if np.linalg.matrix_rank(A, tol=1e-6 * np.abs(A).max()) < 4:
            rec[v] = g[es].mean(0)
        else:
            rec[v] = np.linalg.lstsq(A, g[es], rcond=None)[0][0]
    sq = ((rec[cn] - raw[cn]) ** 2).sum(axis=(0, 2))
    return np.sqrt((vol * sq / 4).sum()), np.sqrt(sq / 4) * np.cbrt(vol), raw.ravel(), rec.ravel()


def test_zz_estimate_matches_numpy_restatement(O):
    import flexmesh as fm

    cfg = fm.configs.make(2, dims=(3, 3, 2))
    rng = np.random.default_rng(4)
    Xn = cfg["X"].reshape(3, -1)
    # a smooth nonlinear field plus noise: the recovered and raw values differ
    u = np.stack([0.05 * Xn[0] ** 2, 0.03 * Xn[0] * Xn[1], -0.02 * Xn[2] ** 2]).T.ravel()
    u += rng.uniform(-1e-3, 1e-3, u.size)
    go, eo, ro, co = O.zz_estimate(cfg["X"], cfg["conn"], u)
    gn, en, rn, cnum = zz_numpy(cfg["X"], cfg["conn"], u)
    scale = np.abs(ro).max()
    assert np.abs(ro - rn).max() < 1e-13 * scale
    assert np.abs(co - cnum).max() < 1e-10 * scale
    assert np.abs(eo - en).max() < 1e-10 * en.max() and abs(go - gn) < 1e-10 * gn
    assert gn > 0


def test_zz_estimate_vanishes_for_homogeneous_and_recovers_linear_fields(O):
    import flexmesh as fm

    cfg = fm.configs.make(2, dims=(3, 2, 2))
    Xn = cfg["X"].reshape(3, -1)
    G = np.array([[0.1, 0.02, -0.03], [0.01, -0.05, 0.04], [0.0, 0.03, 0.08]])
    g, e, raw, rec = O.zz_estimate(cfg["X"], cfg["conn"], (G @ Xn).T.ravel())
    assert abs(g) < 1e-14 and np.abs(e).max() < 1e-14
    assert np.abs(rec.reshape(-1, 3) - np.diag(G)).max() < 1e-13
    assert np.abs(raw.reshape(-1, 3) - np.diag(G)).max() < 1e-14
    # u_x = a x^2 / 2 on an undistorted grid: F11 - 1 is linear in the centroid x, so
    # the SPR fit at an interior node returns a x_node exactly (an evaluation at the
    # origin instead of the node would return ~0)
    cfg0 = fm.configs.make(1, dims=(4, 4, 4))
    X0 = cfg0["X"].reshape(3, -1)
    a = 0.2
    u = np.zeros((X0.shape[1], 3))
    u[:, 0] = 0.5 * a * X0[0] ** 2
    _, _, _, rec = O.zz_estimate(cfg0["X"], cfg0["conn"], u.ravel())
    interior = np.all((X0 > 1e-9) & (X0 < 1 - 1e-9), axis=0)
    assert interior.sum() == 27
    assert np.abs(rec.reshape(-1, 3)[interior, 0] - a * X0[0, interior]).max() < 1e-13
