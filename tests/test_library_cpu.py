"""Host-side checks of the product library that need no GPU: it loads, exports
every entry point the C header declares, and its synthetic-input generators
(mesher, jitter, lognormal fields, boundary data) are exact and reproducible."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "flexmesh", "fm_api.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fm_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import flexmesh

    L = flexmesh.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert b"gfx950" in L.fm_version()


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 2, 4), (8, 8, 8), (16, 2, 2)])
def test_mesher_matches_oracle_bit_exact(O, dims):
    import flexmesh

    X, c = flexmesh.mesh_box(*dims, 8.0, 1.0, 1.0)
    Xo, co = O.box_mesh(*dims, 8.0, 1.0, 1.0)
    assert np.array_equal(X, Xo) and np.array_equal(c, co)


def test_fields_are_seeded_and_lognormal():
    import flexmesh

    a = flexmesh.field_lognormal(200000, 1.0, 0.3, 42)
    b = flexmesh.field_lognormal(200000, 1.0, 0.3, 42)
    c = flexmesh.field_lognormal(200000, 1.0, 0.3, 43)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    z = np.log(a)
    assert abs(np.median(a) - 1.0) < 0.01 and abs(z.std() - 0.3) < 0.003 and abs(z.mean()) < 0.003


def test_jitter_moves_interior_only_and_keeps_orientation():
    import flexmesh
    from flexmesh import configs

    cfg = configs.make(2, dims=(10, 9, 8))
    X0, _ = flexmesh.mesh_box(10, 9, 8)
    gi, gj, gk = configs.grid_index(10, 9, 8)
    bnd = (gi == 0) | (gi == 10) | (gj == 0) | (gj == 9) | (gk == 0) | (gk == 8)
    nn = X0.size // 3
    d = (cfg["X"] - X0).reshape(3, nn)
    assert not d[:, bnd].any()
    assert np.abs(d).max() <= 0.2 / 10 + 1e-15 and np.abs(d[:, ~bnd]).max() > 0.15 / 10
    P = cfg["X"].reshape(3, nn).T[cfg["conn"].reshape(4, -1).T]
    J = np.stack([P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]], axis=2)
    assert np.linalg.det(J).min() > 0


def test_config_sizes_and_boundary_data():
    from flexmesh import configs

    expect = {1: (729, 3072), 2: (274625, 1572864), 3: (2146689, 12582912), 4: (117649, 663552), 5: (279873, 1572864)}
    for k, (nn, ne) in expect.items():
        nx, ny, nz = configs.DIMS[k]
        assert ((nx + 1) * (ny + 1) * (nz + 1), 6 * nx * ny * nz) == (nn, ne)
    cfg = configs.make(1)
    assert cfg["X"].size == 3 * 729 and cfg["conn"].size == 4 * 3072
    # x=0 clamped (3 DOFs x 81 nodes) + u_x on x=1 (81 nodes)
    assert cfg["dir_dofs"].size == 4 * 81 and np.isclose(cfg["dir_vals"].max(), 0.2)
    beam = configs.make(5, dims=(16, 4, 4))
    f = beam["f_ext"].reshape(-1, 3)
    assert np.isclose(f[:, 2].sum(), -beam["traction"] * 1.0) and not f[:, :2].any()


def test_configuration_inner_solve_options():
    """newton_cfg applies a configuration's inner-solve options (config 5:
    extrapolated start, Galerkin start, polynomial coarse level, lagged
    factor) to its FP32-preconditioned and mixed forms only; an FP64 run
    keeps the reference's settings."""
    from flexmesh import configs
    import flexmesh

    beam = configs.make(5, dims=(16, 4, 4))
    c = configs.newton_cfg(beam)
    assert c.prec == flexmesh.FM_MIXED_PC and c.lift == 2 and c.coarse_degree == 4 and c.coarse_additive == 1
    assert c.recycle == 6 and c.lag_precond == 1 and c.inner_tol == 1e-3 and c.inner_tol_max == 0.1
    d = configs.newton_cfg(beam, prec=flexmesh.FM_FP64)
    dflt = flexmesh.Solver.newton_cfg()
    assert d.coarse_degree == 0 and d.recycle == 0 and d.lift == dflt.lift and d.inner_tol_max == dflt.inner_tol_max
    # keyword overrides win over the options
    e = configs.newton_cfg(beam, coarse_degree=2)
    assert e.coarse_degree == 2 and e.coarse_additive == 1
    # configurations without options are unchanged
    cube = configs.make(3, dims=(4, 4, 4))
    assert "inner_opts" not in cube and configs.newton_cfg(cube).coarse_degree == 0
