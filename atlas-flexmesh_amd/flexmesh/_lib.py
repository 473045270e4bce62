"""ctypes binding of libflexmesh.so (the C ABI declared in include/flexmesh/fm_api.h).

The binding has no fallback: if the HIP library is missing or no GPU is
visible, loading raises.  Nothing here computes on the host beyond argument
marshalling.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FM_LIB_PATH selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("FM_LIB_PATH") or os.path.join(HERE, "libflexmesh.so")

FM_OK, FM_EINVAL, FM_ESTATE, FM_EHIP, FM_ENOCONV, FM_EDEVICE = 0, -1, -2, -3, -4, -5
FM_FP64, FM_MIXED, FM_MIXED_PC = 0, 1, 2
KERNEL_CLASSES = ["spmv", "ilu_fwd", "ilu_bwd", "ilu_factor", "assemble", "residual", "pcg_other", "fields",
                  "galerkin", "coarse"]

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class NewtonCfg(C.Structure):
    _fields_ = [("n_steps", C.c_uint32), ("prec", C.c_int), ("newton_rtol", C.c_double), ("max_newton", C.c_int),
                ("pcg_tol", C.c_double), ("pcg_max", C.c_int), ("inner_tol", C.c_double), ("ls_max", C.c_int),
                ("armijo_c", C.c_double), ("lift", C.c_int), ("inner_max", C.c_int),
                ("ls_mode", C.c_int), ("tol_absolute", C.c_double), ("ls_alpha0", C.c_double),
                ("ls_beta", C.c_double), ("reuse_tangent", C.c_int), ("energy_tol", C.c_double),
                ("record_energy", C.c_int), ("inexact", C.c_int), ("recycle", C.c_int),
                ("inner_tol_max", C.c_double), ("coarse_degree", C.c_int),
                ("coarse_additive", C.c_int), ("lag_precond", C.c_int),
                ("extrap_order", C.c_int), ("slope_start", C.c_int),
                ("factor_steps", C.c_int), ("keep_coarse", C.c_int), ("ritz", C.c_int)]


FIELD_NAMES = ("det_F", "cauchy", "von_mises", "pressure", "max_principal", "energy", "max_stretch", "F")


class ElementFieldsOut(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in FIELD_NAMES]


class NewtonStats(C.Structure):
    _fields_ = [("converged", C.c_int), ("newton_iters", C.c_int), ("pcg_iters", C.c_int), ("assemblies", C.c_int),
                ("residuals", C.c_int), ("factorizations", C.c_int), ("r0", C.c_double), ("rel_res", C.c_double),
                ("seconds", C.c_double), ("pcg_cap_hits", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FlexMeshError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"flexmesh error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libflexmesh.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() (or make -C atlas-flexmesh_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    L.fm_version.restype = C.c_char_p
    L.fm_last_error.restype = C.c_char_p
    L.fm_last_error.argtypes = [_vp]
    L.fm_create.argtypes = [C.c_int, C.POINTER(_vp)]
    L.fm_destroy.argtypes = [_vp]
    L.fm_mesh_box.argtypes = [C.c_int] * 3 + [C.c_double] * 3 + [_f64p, _u32p]
    L.fm_mesh_jitter.argtypes = [C.c_uint32, _f64p, _u8p, C.c_double, C.c_uint64]
    L.fm_field_lognormal.argtypes = [C.c_uint32, C.c_double, C.c_double, C.c_uint64, _f64p]
    L.fm_set_mesh.argtypes = [_vp, C.c_uint32, _f64p, C.c_uint32, _u32p]
    L.fm_set_coords.argtypes = [_vp, C.c_uint32, _f64p]
    L.fm_set_material.argtypes = [_vp, _f64p, _f64p]
    L.fm_set_bcs.argtypes = [_vp, C.c_uint32, _vp, _vp]
    L.fm_set_load.argtypes = [_vp, _vp]
    L.fm_build_pattern.argtypes = [_vp, _vp, _vp, C.POINTER(C.c_uint64)]
    L.fm_color.argtypes = [_vp, _vp, C.POINTER(C.c_uint32)]
    L.fm_assemble.argtypes = [_vp, _vp, _vp, _vp, C.c_int]
    L.fm_residual.argtypes = [_vp, _vp, C.c_double, _vp, C.POINTER(C.c_double)]
    L.fm_assemble_residual_only.argtypes = [_vp, _vp, _f64p]
    L.fm_residual_noalloc.argtypes = [_vp, _vp, C.c_double, _vp, C.POINTER(C.c_double)]
    L.fm_set_matrix.argtypes = [_vp, C.c_uint32, _u32p, _u32p, _f64p]
    L.fm_use_tangent.argtypes = [_vp, C.c_int]
    L.fm_spmv.argtypes = [_vp, C.c_int, _f64p, _f64p]
    L.fm_ilu0_factor.argtypes = [_vp, C.c_int, _vp]
    L.fm_ilu0_apply.argtypes = [_vp, C.c_int, _f64p, _f64p]
    L.fm_pcg.argtypes = [_vp, C.c_int, _f64p, _f64p, C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.fm_newton_cfg_default.argtypes = [C.POINTER(NewtonCfg)]
    L.fm_newton_cfg_default.restype = None
    L.fm_newton_step.argtypes = [_vp, C.c_uint32, _vp, C.POINTER(NewtonCfg), C.POINTER(NewtonStats)]
    L.fm_newton_solve_to.argtypes = [_vp, C.c_double, _vp, C.POINTER(NewtonCfg), C.POINTER(NewtonStats)]
    L.fm_newton_history.argtypes = [_vp, C.c_uint32, _vp, _vp, _vp, C.POINTER(C.c_uint32)]
    L.fm_coarse_moments.argtypes = [_vp, C.c_int, C.c_int, _f64p, _f64p]
    L.fm_pcg_history.argtypes = [_vp, C.c_uint32, _vp, _vp, C.POINTER(C.c_uint32)]
    L.fm_get_u.argtypes = [_vp, _f64p]
    L.fm_set_u.argtypes = [_vp, _vp]
    L.fm_adjoint.argtypes = [_vp, _f64p, C.c_double, C.c_int, _vp, _vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.fm_rot_logexp_batch.argtypes = [_vp, C.c_uint32, _f64p, C.c_double, _vp, _vp, _vp]
    L.fm_element_rotations.argtypes = [_vp, _f64p]
    L.fm_element_fields.argtypes = [_vp, _vp, C.POINTER(ElementFieldsOut)]
    L.fm_zz_estimate.argtypes = [_vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_double)]
    L.fm_checkpoint_save.argtypes = [_vp, C.c_char_p, C.c_uint32]
    L.fm_checkpoint_load.argtypes = [_vp, C.c_char_p, C.POINTER(C.c_uint32)]
    L.fm_profile_enable.argtypes = [_vp, C.c_int]
    L.fm_profile_reset.argtypes = [_vp]
    L.fm_profile_get.argtypes = [_vp, _f64p, np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")]
    L.fm_info.argtypes = [_vp, np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")]
    L.fm_sweep_info.argtypes = [_vp, np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")]
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def mesh_box(nx, ny, nz, Lx=1.0, Ly=1.0, Lz=1.0):
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    ne = 6 * nx * ny * nz
    X = np.zeros(3 * nn)
    conn = np.zeros(4 * ne, dtype=np.uint32)
    rc = lib().fm_mesh_box(nx, ny, nz, Lx, Ly, Lz, X, conn)
    if rc != FM_OK:
        raise FlexMeshError(rc, "fm_mesh_box")
    return X, conn


def mesh_jitter(X, interior, amp, seed):
    X = np.array(X, dtype=np.float64)
    rc = lib().fm_mesh_jitter(X.size // 3, X, np.ascontiguousarray(interior, dtype=np.uint8), amp, seed)
    if rc != FM_OK:
        raise FlexMeshError(rc, "fm_mesh_jitter")
    return X


def field_lognormal(n, median, sigma, seed):
    out = np.zeros(n)
    rc = lib().fm_field_lognormal(n, median, sigma, seed, out)
    if rc != FM_OK:
        raise FlexMeshError(rc, "fm_field_lognormal")
    return out


class Solver:
    """One fm_ctx: device-resident mesh, pattern, tangent, factors and state."""

    def __init__(self, device=0):
        self.L = lib()
        h = _vp()
        rc = self.L.fm_create(device, C.byref(h))
        if rc != FM_OK:
            raise FlexMeshError(rc, "fm_create failed (no HIP device visible?)")
        self.h = h
        self.nn = self.ne = 0

    def close(self):
        if self.h:
            self.L.fm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _sizes(self):
        """Node and element counts of the context (the library's, which a
        checkpoint load into an empty context sets)."""
        o = np.zeros(5, dtype=np.uint64)
        self._chk(self.L.fm_info(self.h, o), "fm_info")
        self.nn, self.ne = int(o[3]), int(o[2])
        return self.nn, self.ne

    def _nodal(self, a, what="u"):
        """a as a contiguous FP64 array of 3 nn values (None stays None); the
        library reads exactly that many."""
        if a is None:
            return None
        a = f64(a)
        nn, _ = self._sizes()
        if a.size != 3 * nn:
            raise ValueError(f"{what} has {a.size} values, the mesh has 3 x {nn}")
        return a

    def _nodal_inout(self, u):
        """u passed in and written back: it must already be a contiguous FP64
        array of 3 nn values."""
        if u is None:
            return None
        nn, _ = self._sizes()
        if not (isinstance(u, np.ndarray) and u.dtype == np.float64 and u.flags.c_contiguous and u.size == 3 * nn):
            raise ValueError(f"u must be a contiguous float64 array of 3 x {nn} values")
        return u

    def _chk(self, rc, what):
        if rc != FM_OK:
            msg = self.L.fm_last_error(self.h).decode()
            raise FlexMeshError(rc, f"{what}: {msg}")

    # ---- inputs ---------------------------------------------------------------
    def set_mesh(self, X, conn):
        self.X = f64(X)
        self.conn = u32(conn)
        self.nn, self.ne = self.X.size // 3, self.conn.size // 4
        self._chk(self.L.fm_set_mesh(self.h, self.nn, self.X, self.ne, self.conn), "fm_set_mesh")

    def set_coords(self, X):
        """New node coordinates, same topology (pattern, colouring, schedules kept)."""
        self.X = f64(X)
        self._chk(self.L.fm_set_coords(self.h, self.X.size // 3, self.X), "fm_set_coords")

    def set_material(self, mu, lam):
        self._chk(self.L.fm_set_material(self.h, f64(mu), f64(lam)), "fm_set_material")

    def set_bcs(self, dofs, vals=None):
        d = u32(dofs)
        v = None if vals is None else f64(vals)
        self._keep_bc = (d, v)
        self._chk(self.L.fm_set_bcs(self.h, d.size, _ptr(d), _ptr(v)), "fm_set_bcs")

    def set_load(self, f_ext):
        f = None if f_ext is None else f64(f_ext)
        self._chk(self.L.fm_set_load(self.h, _ptr(f)), "fm_set_load")

    # ---- pattern / colouring ------------------------------------------------
    def pattern(self):
        nnz = C.c_uint64(0)
        self._chk(self.L.fm_build_pattern(self.h, None, None, C.byref(nnz)), "fm_build_pattern")
        self._sizes()
        rp = np.zeros(3 * self.nn + 1, dtype=np.uint32)
        ci = np.zeros(nnz.value, dtype=np.uint32)
        self._chk(self.L.fm_build_pattern(self.h, rp.ctypes.data, ci.ctypes.data, C.byref(nnz)), "fm_build_pattern")
        return rp, ci

    def colors(self):
        self._sizes()
        c = np.zeros(self.ne, dtype=np.uint32)
        nc = C.c_uint32(0)
        self._chk(self.L.fm_color(self.h, c.ctypes.data, C.byref(nc)), "fm_color")
        return c, int(nc.value)

    # ---- assembly ---------------------------------------------------------------
    def assemble(self, u=None, prec=FM_FP64, want_K=True):
        uu = self._nodal(u)
        f = np.zeros(3 * self.nn)
        K = None
        if want_K:
            nnz = C.c_uint64(0)
            self._chk(self.L.fm_build_pattern(self.h, None, None, C.byref(nnz)), "fm_build_pattern")
            K = np.zeros(nnz.value)
        self._chk(self.L.fm_assemble(self.h, _ptr(uu), f.ctypes.data, _ptr(K), prec), "fm_assemble")
        return f, K

    def internal_force(self, u=None):
        """f_int(u) alone (assemble_global_residual_only), no state change."""
        uu = self._nodal(u)
        f = np.zeros(3 * self.nn)
        self._chk(self.L.fm_assemble_residual_only(self.h, _ptr(uu), f), "fm_assemble_residual_only")
        return f

    def residual(self, u=None, load=1.0):
        uu = self._nodal(u)
        R = np.zeros(3 * self.nn)
        nr = C.c_double(0.0)
        self._chk(self.L.fm_residual(self.h, _ptr(uu), load, R.ctypes.data, C.byref(nr)), "fm_residual")
        return R, nr.value

    def residual_noalloc(self, u=None, load=1.0):
        """residual() through the allocation-free entry point (static scratch)."""
        uu = self._nodal(u)
        R = np.zeros(3 * self.nn)
        nr = C.c_double(0.0)
        self._chk(self.L.fm_residual_noalloc(self.h, _ptr(uu), load, R.ctypes.data, C.byref(nr)),
                  "fm_residual_noalloc")
        return R, nr.value

    # ---- sparse linear algebra -------------------------------------------------
    def set_matrix(self, row_ptr, col_idx, vals):
        rp, ci, v = u32(row_ptr), u32(col_idx), f64(vals)
        self._mat = (rp, ci, v)
        self._chk(self.L.fm_set_matrix(self.h, rp.size - 1, rp, ci, v), "fm_set_matrix")
        self._nrows = rp.size - 1
        self._nnz = int(rp[-1])

    def use_tangent(self, eliminate=True):
        self._chk(self.L.fm_use_tangent(self.h, int(eliminate)), "fm_use_tangent")
        info = self.info()
        self._nrows, self._nnz = int(info["n"]), int(info["nnz"])

    def spmv(self, x, prec=FM_FP64):
        y = np.zeros(self._nrows)
        self._chk(self.L.fm_spmv(self.h, prec, f64(x), y), "fm_spmv")
        return y

    def ilu0_factor(self, prec=FM_FP64):
        LU = np.zeros(self._nnz)
        self._chk(self.L.fm_ilu0_factor(self.h, prec, LU.ctypes.data), "fm_ilu0_factor")
        return LU

    def ilu0_apply(self, r, prec=FM_FP64):
        z = np.zeros(self._nrows)
        self._chk(self.L.fm_ilu0_apply(self.h, prec, f64(r), z), "fm_ilu0_apply")
        return z

    def pcg(self, b, x0=None, tol=1e-10, max_it=1000, prec=FM_FP64):
        x = np.zeros(self._nrows) if x0 is None else np.array(x0, dtype=np.float64)
        its = C.c_int(0)
        rel = C.c_double(0.0)
        self._chk(self.L.fm_pcg(self.h, prec, f64(b), x, tol, max_it, C.byref(its), C.byref(rel)), "fm_pcg")
        return x, its.value, rel.value

    # ---- Newton ----------------------------------------------------------------
    @staticmethod
    def newton_cfg(**kw):
        cfg = NewtonCfg()
        lib().fm_newton_cfg_default(C.byref(cfg))
        for k, v in kw.items():
            setattr(cfg, k, v)
        return cfg

    def newton_step(self, step, cfg, u=None, check=True):
        st = NewtonStats()
        uu = self._nodal_inout(u)
        rc = self.L.fm_newton_step(self.h, step, _ptr(uu), C.byref(cfg), C.byref(st))
        if check:
            self._chk(rc, "fm_newton_step")
        return rc, st.as_dict()

    def newton_solve_to(self, load_factor, cfg, u=None, check=True):
        st = NewtonStats()
        rc = self.L.fm_newton_solve_to(self.h, load_factor, _ptr(self._nodal_inout(u)), C.byref(cfg), C.byref(st))
        if check:
            self._chk(rc, "fm_newton_solve_to")
        return rc, st.as_dict()

    def get_u(self):
        nn, _ = self._sizes()
        u = np.zeros(3 * nn)
        self._chk(self.L.fm_get_u(self.h, u), "fm_get_u")
        return u

    def set_u(self, u):
        """Device state u (3 nn); None returns to the unloaded configuration
        (u = 0 at load factor 0, predictor history restarted there)."""
        uu = self._nodal(u)
        self._chk(self.L.fm_set_u(self.h, _ptr(uu)), "fm_set_u")

    def newton_history(self):
        """(residual norms, step sizes, energies) of the last Newton solve."""
        n = C.c_uint32(0)
        self._chk(self.L.fm_newton_history(self.h, 0, None, None, None, C.byref(n)), "fm_newton_history")
        res, alpha, en = np.zeros(n.value), np.zeros(max(n.value - 1, 0)), np.zeros(n.value)
        self._chk(self.L.fm_newton_history(self.h, n.value, res.ctypes.data, alpha.ctypes.data if alpha.size else None,
                                           en.ctypes.data, C.byref(n)), "fm_newton_history")
        return res, alpha, en

    def pcg_history(self):
        """(alpha, beta) of every iteration of the last PCG solve (Lanczos
        coefficients of the preconditioned operator)."""
        n = C.c_uint32(0)
        self._chk(self.L.fm_pcg_history(self.h, 0, None, None, C.byref(n)), "fm_pcg_history")
        a, b = np.zeros(n.value), np.zeros(n.value)
        if n.value:
            self._chk(self.L.fm_pcg_history(self.h, n.value, a.ctypes.data, b.ctypes.data, C.byref(n)),
                      "fm_pcg_history")
        return a, b

    def coarse_moments(self, r, degree, prec):
        """Z^T r over the polynomial fields of total degree <= degree (the
        coarse level's moment kernel in precision prec), 3 nmono values."""
        rr = self._nodal(r, "r")
        out = np.zeros(3 * (4, 10, 20, 35)[degree - 1])
        self._chk(self.L.fm_coarse_moments(self.h, prec, degree, rr, out), "fm_coarse_moments")
        return out

    # ---- adjoint / rotations / checkpoint --------------------------------------
    def adjoint(self, dLdu, tol=1e-12, max_it=5000):
        g = self._nodal(dLdu, "dL/du")
        dmu = np.zeros(self.ne)
        dlam = np.zeros(self.ne)
        dl = C.c_double(0.0)
        its = C.c_int(0)
        self._chk(self.L.fm_adjoint(self.h, g, tol, max_it, dmu.ctypes.data, dlam.ctypes.data, C.byref(dl),
                                    C.byref(its)), "fm_adjoint")
        return dmu, dlam, dl.value, its.value

    def rot_logexp(self, F, alpha=1.0):
        F = f64(F).reshape(-1, 9)
        n = F.shape[0]
        R = np.zeros((n, 9))
        w = np.zeros((n, 3))
        Q = np.zeros((n, 9))
        self._chk(self.L.fm_rot_logexp_batch(self.h, n, F.ravel(), alpha, R.ctypes.data, w.ctypes.data,
                                             Q.ctypes.data), "fm_rot_logexp_batch")
        return R.reshape(n, 3, 3), w, Q.reshape(n, 3, 3)

    def element_rotations(self):
        self._sizes()
        w = np.zeros(3 * self.ne)
        self._chk(self.L.fm_element_rotations(self.h, w), "fm_element_rotations")
        return w.reshape(self.ne, 3)

    def element_fields(self, u=None, fields=FIELD_NAMES):
        """Per-element fields of u (None: the device state), element order:
        cauchy is 6 x ne (xx, yy, zz, yz, xz, xy), F is 9 x ne (F_ij row-major)."""
        uu = self._nodal(u)
        out = {}
        o = ElementFieldsOut()
        for k in fields:
            out[k] = np.zeros({"cauchy": (6, self.ne), "F": (9, self.ne)}.get(k, (self.ne,)))
            setattr(o, k, out[k].ctypes.data)
        self._chk(self.L.fm_element_fields(self.h, _ptr(uu), C.byref(o)), "fm_element_fields")
        return out

    def zz_estimate(self, u=None):
        """ZZ estimate of u (None: the device state): (global, element indicators,
        node averages 3*nn, recovered values 3*nn)."""
        uu = self._nodal(u)
        err, raw, rec = np.zeros(self.ne), np.zeros(3 * self.nn), np.zeros(3 * self.nn)
        g = C.c_double(0.0)
        self._chk(self.L.fm_zz_estimate(self.h, _ptr(uu), err.ctypes.data, raw.ctypes.data, rec.ctypes.data,
                                        C.byref(g)), "fm_zz_estimate")
        return g.value, err, raw, rec

    def checkpoint_save(self, path, load_step):
        self._chk(self.L.fm_checkpoint_save(self.h, path.encode(), load_step), "fm_checkpoint_save")

    def checkpoint_load(self, path):
        s = C.c_uint32(0)
        self._chk(self.L.fm_checkpoint_load(self.h, path.encode(), C.byref(s)), "fm_checkpoint_load")
        self._sizes()  # an empty context has taken the file's mesh
        return s.value

    # ---- measurement ------------------------------------------------------------
    def profile(self, on=True):
        self._chk(self.L.fm_profile_enable(self.h, int(on)), "fm_profile_enable")

    def profile_reset(self):
        self._chk(self.L.fm_profile_reset(self.h), "fm_profile_reset")

    def profile_get(self):
        ms = np.zeros(len(KERNEL_CLASSES))
        n = np.zeros(len(KERNEL_CLASSES), dtype=np.uint64)
        self._chk(self.L.fm_profile_get(self.h, ms, n), "fm_profile_get")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(KERNEL_CLASSES)}

    def info(self):
        o = np.zeros(5, dtype=np.uint64)
        self._chk(self.L.fm_info(self.h, o), "fm_info")
        return dict(zip(["n", "nnz", "ne", "nn", "ncolors"], (int(v) for v in o)))

    def sweep_info(self):
        o = np.zeros(10, dtype=np.uint64)
        self._chk(self.L.fm_sweep_info(self.h, o), "fm_sweep_info")
        d = dict(zip(["chain", "levels_fwd", "levels_bwd", "chains", "bundles", "wave_steps", "lanes_per_node",
                      "gj", "gk", "node_spmv"], (int(v) for v in o)))
        d["sweep"] = "chain" if d.pop("chain") else "level"
        d["spmv"] = "node" if d.pop("node_spmv") else "csr"
        return d
