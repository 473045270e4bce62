"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against and the
CPU baseline bench.py times beside it.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class Problem(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32), ("n_elems", C.c_uint32),
        ("X", C.c_void_p), ("conn", C.c_void_p), ("mu", C.c_void_p), ("lam", C.c_void_p),
        ("n_dir", C.c_uint32), ("dir_dofs", C.c_void_p), ("dir_vals", C.c_void_p),
        ("f_ext", C.c_void_p),
        ("n_steps", C.c_uint32),
        ("newton_rtol", C.c_double), ("max_newton", C.c_int),
        ("pcg_tol", C.c_double), ("pcg_max", C.c_int),
        ("ls_max", C.c_int), ("armijo_c", C.c_double), ("lift", C.c_int),
        ("colors", C.c_void_p), ("n_colors", C.c_uint32),
        ("row_ptr", C.c_void_p), ("col_idx", C.c_void_p), ("ls_mode", C.c_int),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("converged", C.c_int), ("newton_iters", C.c_int), ("pcg_iters", C.c_int),
        ("assemblies", C.c_int), ("residuals", C.c_int),
        ("r0", C.c_double), ("rel_res", C.c_double),
        ("t_assemble", C.c_double), ("t_residual", C.c_double), ("t_factor", C.c_double),
        ("t_pcg", C.c_double), ("t_total", C.c_double), ("factorizations", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_box_mesh.argtypes = [C.c_int] * 3 + [C.c_double] * 3 + [_f64p, _u32p]
        L.orc_inv3.argtypes = [_f64p, _f64p]
        L.orc_pk1_refform.argtypes = [_f64p, C.c_double, C.c_double, _f64p]
        L.orc_pk1.argtypes = [_f64p, C.c_double, C.c_double, _f64p]
        L.orc_energy.argtypes = [_f64p, C.c_double, C.c_double]
        L.orc_energy.restype = C.c_double
        L.orc_element.argtypes = [_f64p, _f64p, C.c_double, C.c_double, _f64p, _f64p]
        L.orc_element.restype = C.c_int
        L.orc_pattern_rows.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p]
        L.orc_pattern_rows.restype = C.c_uint64
        L.orc_pattern_cols.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p, _u32p]
        L.orc_color.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p]
        L.orc_color.restype = C.c_uint32
        L.orc_assemble.argtypes = [C.c_uint32, C.c_uint32, _f64p, _u32p, _f64p, _f64p, _u32p, C.c_uint32,
                                   _f64p, _u32p, _u32p, C.c_void_p, _f64p]
        L.orc_eliminate.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, _u8p]
        L.orc_spmv.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, _f64p, _f64p]
        L.orc_ilu0.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, _f64p, _f64p]
        L.orc_ilu0_apply.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, _f64p, _f64p, _f64p]
        L.orc_pcg.argtypes = [C.c_uint32, _u32p, _u32p, _f64p, _f64p, _f64p, C.c_double, C.c_int,
                              C.POINTER(C.c_double)]
        L.orc_pcg.restype = C.c_int
        L.orc_newton_step.argtypes = [C.POINTER(Problem), C.c_uint32, _f64p, C.POINTER(StepStats)]
        L.orc_newton_step.restype = C.c_int
        L.orc_sym_eig3.argtypes = [_f64p, _f64p]
        L.orc_cauchy.argtypes = [_f64p, C.c_double, C.c_double, _f64p]
        L.orc_von_mises.argtypes = [_f64p]
        L.orc_von_mises.restype = C.c_double
        L.orc_element_fields.argtypes = [C.c_uint32, _f64p, C.c_uint32, _u32p, _f64p, _f64p, _vp] + [_vp] * 8
        L.orc_zz_estimate.argtypes = [C.c_uint32, _f64p, C.c_uint32, _u32p, _vp, _vp, _vp, _vp]
        L.orc_zz_estimate.restype = C.c_double
        L.orc_polar.argtypes = [_f64p, _f64p]
        L.orc_polar.restype = C.c_int
        L.orc_rot_log.argtypes = [_f64p, _f64p]
        L.orc_rot_exp.argtypes = [_f64p, _f64p]
        L.orc_adjoint.argtypes = [C.POINTER(Problem), _f64p, _f64p, C.c_double, C.c_int, _f64p, _f64p,
                                  C.POINTER(C.c_double), C.c_void_p]
        L.orc_adjoint.restype = C.c_int
        L.orc_total_energy.argtypes = [C.POINTER(Problem), C.c_double, _f64p]
        L.orc_total_energy.restype = C.c_double
        _lib = L
    return _lib


def set_threads(n):
    lib().orc_set_threads(int(n))


def box_mesh(nx, ny, nz, Lx=1.0, Ly=1.0, Lz=1.0):
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    ne = 6 * nx * ny * nz
    X = np.zeros(3 * nn)
    conn = np.zeros(4 * ne, dtype=np.uint32)
    lib().orc_box_mesh(nx, ny, nz, Lx, Ly, Lz, X, conn)
    return X, conn


def pk1_refform(F, mu, lam):
    P = np.zeros(9)
    lib().orc_pk1_refform(np.ascontiguousarray(F, dtype=np.float64).ravel(), mu, lam, P)
    return P


def pk1(F, mu, lam):
    P = np.zeros(9)
    lib().orc_pk1(np.ascontiguousarray(F, dtype=np.float64).ravel(), mu, lam, P)
    return P


def energy(F, mu, lam):
    return lib().orc_energy(np.ascontiguousarray(F, dtype=np.float64).ravel(), mu, lam)


def element(Xe, ue, mu, lam):
    f = np.zeros(12)
    K = np.zeros(144)
    skipped = lib().orc_element(np.ascontiguousarray(Xe, dtype=np.float64).ravel(),
                                np.ascontiguousarray(ue, dtype=np.float64).ravel(), mu, lam, f, K)
    return f, K.reshape(12, 12), skipped


def pattern(nn, conn):
    ne = conn.size // 4
    row_ptr = np.zeros(3 * nn + 1, dtype=np.uint32)
    nnz = lib().orc_pattern_rows(nn, ne, conn, row_ptr)
    col_idx = np.zeros(nnz, dtype=np.uint32)
    lib().orc_pattern_cols(nn, ne, conn, row_ptr, col_idx)
    return row_ptr, col_idx


def color(nn, conn):
    ne = conn.size // 4
    colors = np.zeros(ne, dtype=np.uint32)
    nc = lib().orc_color(nn, ne, conn, colors)
    return colors, int(nc)


def assemble(X, conn, mu, lam, colors, n_colors, u, row_ptr, col_idx, want_K=True):
    nn = X.size // 3
    ne = conn.size // 4
    f = np.zeros(3 * nn)
    K = np.zeros(int(row_ptr[-1])) if want_K else None
    lib().orc_assemble(nn, ne, X, conn, mu, lam, colors, n_colors, u, row_ptr, col_idx,
                       K.ctypes.data if want_K else None, f)
    return K, f


def eliminate(row_ptr, col_idx, K, is_dir):
    lib().orc_eliminate(row_ptr.size - 1, row_ptr, col_idx, K, is_dir.astype(np.uint8))


def spmv(row_ptr, col_idx, val, x):
    y = np.zeros(row_ptr.size - 1)
    lib().orc_spmv(row_ptr.size - 1, row_ptr, col_idx, val, np.ascontiguousarray(x, dtype=np.float64), y)
    return y


def ilu0(row_ptr, col_idx, val):
    n = row_ptr.size - 1
    LU = np.zeros(val.size)
    diag = np.zeros(n)
    lib().orc_ilu0(n, row_ptr, col_idx, val, LU, diag)
    return LU, diag


def ilu0_apply(row_ptr, col_idx, LU, diag, x):
    y = np.zeros(row_ptr.size - 1)
    lib().orc_ilu0_apply(row_ptr.size - 1, row_ptr, col_idx, LU, diag, np.ascontiguousarray(x, dtype=np.float64), y)
    return y


def pcg(row_ptr, col_idx, val, b, tol=1e-10, max_iter=1000, x0=None):
    n = row_ptr.size - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    rel = C.c_double(0.0)
    its = lib().orc_pcg(n, row_ptr, col_idx, val, np.ascontiguousarray(b, dtype=np.float64), x, tol, max_iter,
                        C.byref(rel))
    return x, its, rel.value


def sym_eig3(A):
    w = np.zeros(3)
    lib().orc_sym_eig3(np.ascontiguousarray(A, dtype=np.float64).ravel(), w)
    return w


def cauchy(F, mu, lam):
    s = np.zeros(9)
    lib().orc_cauchy(np.ascontiguousarray(F, dtype=np.float64).ravel(), mu, lam, s)
    return s.reshape(3, 3)


def von_mises(sigma):
    return lib().orc_von_mises(np.ascontiguousarray(sigma, dtype=np.float64).ravel())


FIELD_NAMES = ("det_F", "cauchy", "von_mises", "pressure", "max_principal", "energy", "max_stretch", "F")


def element_fields(X, conn, mu, lam, u):
    """Per-element fields in element order: cauchy is 6 x ne (xx, yy, zz, yz, xz, xy), F is 9 x ne."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.uint32)
    nn, ne = X.size // 3, conn.size // 4
    u = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
    out = {k: np.zeros(ne) for k in FIELD_NAMES}
    out["cauchy"] = np.zeros((6, ne))
    out["F"] = np.zeros((9, ne))
    lib().orc_element_fields(nn, X, ne, conn, np.ascontiguousarray(mu, dtype=np.float64),
                             np.ascontiguousarray(lam, dtype=np.float64), None if u is None else u.ctypes.data,
                             *[out[k].ctypes.data for k in FIELD_NAMES])
    return out


def zz_estimate(X, conn, u):
    """ZZ estimate: (global, element indicators, node averages 3*nn, recovered 3*nn)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.uint32)
    nn, ne = X.size // 3, conn.size // 4
    u = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
    err, raw, rec = np.zeros(ne), np.zeros(3 * nn), np.zeros(3 * nn)
    g = lib().orc_zz_estimate(nn, X, ne, conn, None if u is None else u.ctypes.data, err.ctypes.data,
                              raw.ctypes.data, rec.ctypes.data)
    return g, err, raw, rec


def polar(F):
    R = np.zeros(9)
    st = lib().orc_polar(np.ascontiguousarray(F, dtype=np.float64).ravel(), R)
    return R.reshape(3, 3), st


def rot_log(R):
    w = np.zeros(3)
    lib().orc_rot_log(np.ascontiguousarray(R, dtype=np.float64).ravel(), w)
    return w


def rot_exp(w):
    R = np.zeros(9)
    lib().orc_rot_exp(np.ascontiguousarray(w, dtype=np.float64).ravel(), R)
    return R.reshape(3, 3)


class NewtonProblem:
    """Holds the arrays an orc_problem points at (keeps them alive)."""

    def __init__(self, X, conn, mu, lam, dir_dofs, dir_vals, f_ext, n_steps, newton_rtol=1e-8,
                 max_newton=30, pcg_tol=1e-10, pcg_max=300, ls_max=10, armijo_c=1e-4, lift=1,
                 colors=None, n_colors=None, row_ptr=None, col_idx=None, ls_mode=0):
        self.X = np.ascontiguousarray(X, dtype=np.float64)
        self.conn = np.ascontiguousarray(conn, dtype=np.uint32)
        self.mu = np.ascontiguousarray(mu, dtype=np.float64)
        self.lam = np.ascontiguousarray(lam, dtype=np.float64)
        self.dir_dofs = np.ascontiguousarray(dir_dofs, dtype=np.uint32)
        self.dir_vals = np.ascontiguousarray(dir_vals, dtype=np.float64)
        self.f_ext = None if f_ext is None else np.ascontiguousarray(f_ext, dtype=np.float64)
        nn = self.X.size // 3
        if row_ptr is None:
            row_ptr, col_idx = pattern(nn, self.conn)
        if colors is None:
            colors, n_colors = color(nn, self.conn)
        self.row_ptr, self.col_idx = row_ptr, col_idx
        self.colors, self.n_colors = colors, n_colors
        p = Problem()
        p.n_nodes = nn
        p.n_elems = self.conn.size // 4
        p.X = self.X.ctypes.data
        p.conn = self.conn.ctypes.data
        p.mu = self.mu.ctypes.data
        p.lam = self.lam.ctypes.data
        p.n_dir = self.dir_dofs.size
        p.dir_dofs = self.dir_dofs.ctypes.data
        p.dir_vals = self.dir_vals.ctypes.data
        p.f_ext = None if self.f_ext is None else self.f_ext.ctypes.data
        p.n_steps = n_steps
        p.newton_rtol = newton_rtol
        p.max_newton = max_newton
        p.pcg_tol = pcg_tol
        p.pcg_max = pcg_max
        p.ls_max = ls_max
        p.armijo_c = armijo_c
        p.lift = lift
        p.colors = self.colors.ctypes.data
        p.n_colors = n_colors
        p.row_ptr = self.row_ptr.ctypes.data
        p.col_idx = self.col_idx.ctypes.data
        p.ls_mode = ls_mode
        self.p = p

    def step(self, s, u):
        st = StepStats()
        rc = lib().orc_newton_step(C.byref(self.p), s, u, C.byref(st))
        return rc, st.as_dict()

    def energy(self, lf, u):
        return lib().orc_total_energy(C.byref(self.p), lf, np.ascontiguousarray(u, dtype=np.float64))

    def adjoint(self, u, dLdu, tol=1e-12, max_iter=5000):
        ne = self.conn.size // 4
        dmu = np.zeros(ne)
        dlam = np.zeros(ne)
        dload = C.c_double(0.0)
        lam_v = np.zeros(u.size)
        its = lib().orc_adjoint(C.byref(self.p), u, np.ascontiguousarray(dLdu, dtype=np.float64), tol, max_iter,
                                dmu, dlam, C.byref(dload), lam_v.ctypes.data)
        return dmu, dlam, dload.value, lam_v, its
