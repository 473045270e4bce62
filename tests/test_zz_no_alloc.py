"""The allocation-free residual entry point (fm_residual_noalloc) against
fm_residual through the C ABI: the same kernels and reduction order, so R and
||R|| must be bit-identical; two calls in a row on one GPU's shared scratch,
the second on a smaller mesh (stale partials of the first call beyond its
range), then the first mesh again; and the size and set-up checks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fm():
    import flexmesh

    flexmesh.lib()
    return flexmesh


def problem(fm, dims, seed):
    cfg = fm.configs.make(3, dims=dims)
    u = np.random.default_rng(seed).uniform(-0.02, 0.02, cfg["X"].size)
    return cfg, u


@pytest.mark.parametrize("k", [3, 5])
def test_noalloc_residual_equals_residual(fm, k):
    S = fm.Solver(0)
    try:
        for dims, seed, load in (((24, 20, 22), 1, 0.7), ((5, 4, 6), 2, 0.3), ((24, 20, 22), 3, 1.0)):
            cfg, u = problem(fm, dims, seed)
            if k == 5:
                # a traction load: the f_ext term of R
                cfg["f_ext"] = np.random.default_rng(seed + 10).standard_normal(cfg["X"].size) * 1e-3
            fm.configs.load(S, cfg)
            R, nr = S.residual(u, load)
            for _ in range(2):
                Rn, nrn = S.residual_noalloc(u, load)
                assert np.array_equal(Rn, R) and nrn == nr
                assert not Rn[cfg["dir_dofs"]].any() and nr > 0
    finally:
        S.close()


def test_noalloc_residual_checks(fm):
    S = fm.Solver(0)
    try:
        cfg, u = problem(fm, (3, 3, 3), 4)
        S.set_mesh(cfg["X"], cfg["conn"])
        # no material yet: refused, nothing made
        with pytest.raises(fm.FlexMeshError) as e:
            S.residual_noalloc(u, 1.0)
        assert e.value.code == fm.FM_ESTATE
        fm.configs.load(S, cfg)
        R, nr = S.residual(u, 1.0)
        Rn, nrn = S.residual_noalloc(u, 1.0)
        assert np.array_equal(Rn, R) and nrn == nr
        # beyond the scratch limit of 8,388,608 DOFs: refused before any work
        nn = 2_796_203
        X = np.zeros(3 * nn)
        X[:4] = (0.0, 1.0, 0.0, 0.0)
        X[nn + 2] = 1.0
        X[2 * nn + 3] = 1.0
        S.set_mesh(X, np.array([0, 1, 2, 3], np.uint32))
        with pytest.raises(fm.FlexMeshError) as e:
            S.residual_noalloc(None, 1.0)
        assert e.value.code == fm.FM_EINVAL and "8388608" in str(e.value)
    finally:
        S.close()
