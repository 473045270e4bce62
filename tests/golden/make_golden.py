"""Writes the golden vectors under tests/golden/ from the compiled reference.

Run in the container that holds the reference (it needs oracle/_ref/libref.so,
built by oracle/build_ref.sh from the reference sources).  Each fixture stores
inputs and the reference's outputs only; tests read the fixtures and never run
this script.

    python tests/golden/make_golden.py              # every fixture
    python tests/golden/make_golden.py checkpoint   # the checkpoint fixture only
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402  (box mesher for the Kuhn inputs)

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def load_ref():
    path = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    if not os.path.exists(path):
        raise SystemExit("oracle/_ref/libref.so missing: run oracle/build_ref.sh first")
    L = C.CDLL(path)
    L.ref_pk1.argtypes = [f64p, C.c_double, C.c_double, f64p]
    L.ref_cube_mesh.argtypes = [C.c_int, f64p, u32p]
    L.ref_cube_mesh.restype = C.c_uint32
    L.ref_csr.argtypes = [C.c_uint32, f64p, C.c_uint32, u32p, u32p, C.c_void_p]
    L.ref_csr.restype = C.c_uint32
    L.ref_color.argtypes = [C.c_uint32, f64p, C.c_uint32, u32p, u32p]
    L.ref_color.restype = C.c_uint32
    L.ref_ilu0.argtypes = [C.c_uint32, u32p, u32p, f64p, f64p, f64p, f64p]
    L.ref_ilu0_solve.argtypes = [C.c_uint32, u32p, u32p, f64p, f64p, f64p]
    L.ref_pcg.argtypes = [C.c_uint32, u32p, u32p, f64p, f64p, f64p, C.c_double, C.c_int]
    L.ref_pcg.restype = C.c_int
    L.ref_expm.argtypes = [f64p, f64p]
    L.ref_logm.argtypes = [f64p, f64p]
    return L


def poisson(grid_n, sort_cols):
    """The 5-point matrix of TESTS/external_validation_runner.cpp:65-124
    (row index i*grid_n + j; diagonal first, then i-1, i+1, j-1, j+1)."""
    n = grid_n * grid_n
    h = 1.0 / (grid_n + 1)
    ih2 = 1.0 / (h * h)
    rp = [0]
    ci, va = [], []
    xs = np.zeros(n)
    for i in range(grid_n):
        for j in range(grid_n):
            row = i * grid_n + j
            xs[row] = np.sin(np.pi * (i + 1) * h) * np.sin(np.pi * (j + 1) * h)
            ents = [(row, 4.0 * ih2)]
            if i > 0:
                ents.append(((i - 1) * grid_n + j, -ih2))
            if i + 1 < grid_n:
                ents.append(((i + 1) * grid_n + j, -ih2))
            if j > 0:
                ents.append((i * grid_n + j - 1, -ih2))
            if j + 1 < grid_n:
                ents.append((i * grid_n + j + 1, -ih2))
            if sort_cols:
                ents.sort()
            for c, v in ents:
                ci.append(c)
                va.append(v)
            rp.append(len(ci))
    rp = np.array(rp, dtype=np.uint32)
    ci = np.array(ci, dtype=np.uint32)
    va = np.array(va)
    # b = A x_exact with the matvec's storage-order sum (fem_types.hpp:310-319)
    b = np.array([sum(va[k] * xs[ci[k]] for k in range(rp[r], rp[r + 1])) for r in range(n)])
    return rp, ci, va, b, xs


def elastic_matrix(N, seed):
    """Tangent of a Kuhn mesh at a seeded displacement (oracle assembly), with
    x=0 clamped by elimination: a small SPD elasticity matrix."""
    X, conn = O.box_mesh(N, N, N)
    nn = X.size // 3
    ne = conn.size // 4
    rp, ci = O.pattern(nn, conn)
    colors, nc = O.color(nn, conn)
    rng = np.random.default_rng(seed)
    u = rng.uniform(-0.02, 0.02, 3 * nn)
    mu = np.exp(0.3 * rng.standard_normal(ne))
    lam = 1.5 * mu
    K, _ = O.assemble(X, conn, mu, lam, colors, nc, u, rp, ci)
    is_dir = np.zeros(3 * nn, dtype=np.uint8)
    x0 = np.where(X[:nn] == 0.0)[0]
    for d in range(3):
        is_dir[3 * x0 + d] = 1
    O.eliminate(rp, ci, K, is_dir)
    return rp, ci, K


def stress_fixture(L):
    """compute_cauchy_stress (FEM/constitutive_models.hpp:116-148) and
    compute_von_mises (FEM/stress_extraction.hpp:143-157) of the compiled
    reference, on the deformation gradients of the pk1 fixture plus a rotation
    and a pure shear."""
    L.ref_cauchy.argtypes = [f64p, C.c_double, C.c_double, f64p]
    L.ref_von_mises.argtypes = [f64p]
    L.ref_von_mises.restype = C.c_double
    rng = np.random.default_rng(20240612)
    Fs = [np.array([[1.1, 0.05, 0.0], [-0.02, 0.95, 0.03], [0.01, 0.0, 1.05]]), np.eye(3)]
    for _ in range(24):
        Fs.append(np.eye(3) + rng.uniform(-0.3, 0.3, (3, 3)))
    Fs.append(np.diag([0.05, 0.1, 0.9]))  # J < 0.01: clamps of ln J and of 1/J
    c, s_ = np.cos(0.7), np.sin(0.7)
    Fs.append(np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]]))  # rotation: sigma = 0
    Fs.append(np.array([[1.0, 0.2, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]))  # simple shear
    mats = [(1.0, 0.3), (1.0, 0.49), (210e9, 0.3)]
    Fa, Ea, nua, Sa, Va = [], [], [], [], []
    for F in Fs:
        for E, nu in mats:
            S = np.zeros(9)
            L.ref_cauchy(np.ascontiguousarray(F.ravel()), E, nu, S)
            Fa.append(F.ravel())
            Ea.append(E)
            nua.append(nu)
            Sa.append(S)
            Va.append(L.ref_von_mises(np.ascontiguousarray(S)))
    np.savez(os.path.join(HERE, "stress.npz"), F=np.array(Fa), E=np.array(Ea), nu=np.array(nua),
             sigma=np.array(Sa), von_mises=np.array(Va))


def checkpoint_fixture(L):
    """CheckpointFile::save bytes for a jittered 2x3x2 Kuhn box with a smooth
    displacement, E = 1, nu = 0.3, load step 7, adapt iteration 0."""
    import tempfile

    L.ref_checkpoint_save.argtypes = [C.c_char_p, C.c_uint32, f64p, f64p, C.c_uint32, u32p, C.c_double,
                                      C.c_double, C.c_uint32, C.c_uint32]
    L.ref_checkpoint_save.restype = C.c_int
    X, conn = O.box_mesh(2, 3, 2)
    nn, ne = X.size // 3, conn.size // 4
    P = X.reshape(3, nn).T
    rng = np.random.default_rng(21)
    P = P + rng.uniform(-0.05, 0.05, P.shape)
    X = np.ascontiguousarray(P.T).ravel()
    u = np.ascontiguousarray(np.stack([0.12 * P[:, 0] + 0.03 * np.sin(3 * P[:, 1]),
                                       -0.04 * P[:, 1] + 0.02 * P[:, 0] * P[:, 2],
                                       -0.05 * P[:, 2] + 0.01 * np.cos(2 * P[:, 0])], axis=1).ravel())
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "ref.ckpt")
        assert L.ref_checkpoint_save(path.encode(), nn, X, u, ne, conn, 1.0, 0.3, 7, 0) == 0
        data = np.frombuffer(open(path, "rb").read(), dtype=np.uint8).copy()
    np.savez_compressed(os.path.join(HERE, "checkpoint.npz"), X=X, conn=conn, u=u, E=np.array([1.0]),
                        nu=np.array([0.3]), load_step=np.array([7], np.uint32), ref_bytes=data)


def main():
    if sys.argv[1:] == ["checkpoint"]:
        checkpoint_fixture(load_ref())
        return
    L = load_ref()
    out = {}

    # -- a5: compute_pk1_stress ------------------------------------------------
    rng = np.random.default_rng(20240611)
    Fs = [np.array([[1.1, 0.05, 0.0], [-0.02, 0.95, 0.03], [0.01, 0.0, 1.05]]), np.eye(3)]
    for _ in range(30):
        Fs.append(np.eye(3) + rng.uniform(-0.3, 0.3, (3, 3)))
    Fs.append(np.diag([0.05, 0.1, 0.9]))  # J < 0.01: ln J clamp
    Fs.append(np.diag([1.0, 1.0, -1.0]) @ (np.eye(3) + 0.1 * rng.standard_normal((3, 3))))  # det < 0
    mats = [(1.0, 0.3), (1.0, 0.45), (1.0, 0.49), (210e9, 0.3)]
    Fa, Ea, nua, Pa = [], [], [], []
    for F in Fs:
        for E, nu in mats:
            P = np.zeros(9)
            L.ref_pk1(np.ascontiguousarray(F.ravel()), E, nu, P)
            Fa.append(F.ravel())
            Ea.append(E)
            nua.append(nu)
            Pa.append(P)
    np.savez(os.path.join(HERE, "pk1.npz"), F=np.array(Fa), E=np.array(Ea), nu=np.array(nua), P=np.array(Pa))

    # -- a2 / a10 / a13: mesher numbering, CSR row structure, coloring ----------
    mesh = {}
    for N in (1, 2, 4, 8):
        nn = (N + 1) ** 3
        Xr = np.zeros(3 * nn)
        connr = np.zeros(4 * 6 * N ** 3, dtype=np.uint32)
        ne = L.ref_cube_mesh(N, Xr, connr)
        mesh[f"ref_X_{N}"] = Xr
        mesh[f"ref_conn_{N}"] = connr
        rp = np.zeros(3 * nn + 1, dtype=np.uint32)
        L.ref_csr(nn, Xr, ne, connr, rp, None)
        mesh[f"ref_rowptr_refmesh_{N}"] = rp
        Xk, connk = O.box_mesh(N, N, N)
        rpk = np.zeros(3 * nn + 1, dtype=np.uint32)
        L.ref_csr(nn, Xk, ne, connk, rpk, None)
        mesh[f"ref_rowptr_kuhn_{N}"] = rpk
        colk = np.zeros(ne, dtype=np.uint32)
        mesh[f"ref_ncolors_kuhn_{N}"] = np.array([L.ref_color(nn, Xk, ne, connk, colk)])
        mesh[f"ref_colors_kuhn_{N}"] = colk
        if N <= 4:
            colr = np.zeros(ne, dtype=np.uint32)
            mesh[f"ref_ncolors_refmesh_{N}"] = np.array([L.ref_color(nn, Xr, ne, connr, colr)])
            mesh[f"ref_colors_refmesh_{N}"] = colr
    np.savez_compressed(os.path.join(HERE, "mesh_pattern_color.npz"), **mesh)

    # -- a14: ILU(0) factors ----------------------------------------------------
    ilu = {}
    for name, (rp, ci, va) in {
        "poisson16_refcols": poisson(16, False)[:3],
        "poisson16_sorted": poisson(16, True)[:3],
        "elastic2": elastic_matrix(2, 7),
    }.items():
        n = rp.size - 1
        Lv = np.zeros(va.size)
        Uv = np.zeros(va.size)
        dg = np.zeros(n)
        L.ref_ilu0(n, rp, ci, va, Lv, Uv, dg)
        x = np.random.default_rng(5).standard_normal(n)
        y = np.zeros(n)
        L.ref_ilu0_solve(n, rp, ci, va, x, y)
        ilu.update({f"{name}_rp": rp, f"{name}_ci": ci, f"{name}_val": va, f"{name}_L": Lv, f"{name}_U": Uv,
                    f"{name}_diag": dg, f"{name}_x": x, f"{name}_y": y})
    np.savez_compressed(os.path.join(HERE, "ilu0.npz"), **ilu)

    # -- a16: pcg_solve_ilu0 iteration counts and solutions -----------------------
    pcg = {}
    for g in (16, 32, 48, 96):
        for sort_cols in (False, True):
            rp, ci, va, b, xs = poisson(g, sort_cols)
            x = np.zeros(rp.size - 1)
            its = L.ref_pcg(rp.size - 1, rp, ci, va, b, x, 1e-10, 4000)
            tag = f"poisson{g}_{'sorted' if sort_cols else 'refcols'}"
            pcg[f"{tag}_iters"] = np.array([its])
            if g <= 48 or sort_cols:
                pcg[f"{tag}_x"] = x
    rp, ci, va = elastic_matrix(2, 11)
    b = np.random.default_rng(3).standard_normal(rp.size - 1)
    x = np.zeros(rp.size - 1)
    its = L.ref_pcg(rp.size - 1, rp, ci, va, b, x, 1e-10, 300)
    pcg.update({"elastic2_rp": rp, "elastic2_ci": ci, "elastic2_val": va, "elastic2_b": b, "elastic2_x": x,
                "elastic2_iters": np.array([its])})
    np.savez_compressed(os.path.join(HERE, "pcg.npz"), **pcg)

    # -- a20: exponential_map / matrix_logarithm on rotations ----------------------
    xi_ref = np.array([[0.10, 0.02, 0.05], [-0.02, 0.15, -0.03], [-0.05, -0.03, -0.25]])
    ws, Es, Ls = [], [], []
    rng = np.random.default_rng(99)
    for ang in (1e-3, 1e-2, 0.1, 0.5, 1.0, np.pi / 3, np.pi / 2):
        for _ in range(3):
            n = rng.standard_normal(3)
            n /= np.linalg.norm(n)
            w = ang * n
            K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
            E = np.zeros(9)
            L.ref_expm(np.ascontiguousarray(K.T.ravel()), E)  # column-major in, column-major out
            Em = E.reshape(3, 3).T
            Lg = np.zeros(9)
            L.ref_logm(np.ascontiguousarray(Em.T.ravel()), Lg)
            ws.append(w)
            Es.append(Em.ravel())
            Ls.append(Lg.reshape(3, 3).T.ravel())
    E = np.zeros(9)
    L.ref_expm(np.ascontiguousarray(xi_ref.T.ravel()), E)
    np.savez(os.path.join(HERE, "lie.npz"), w=np.array(ws), expm=np.array(Es), logm=np.array(Ls),
             xi_ref=xi_ref.ravel(), expm_xi_ref=E.reshape(3, 3).T.ravel())
    stress_fixture(L)
    checkpoint_fixture(L)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
