"""Per-kernel-class HIP-event timing through the C ABI (fm_profile_enable /
fm_profile_get): with sampling (on = k > 1) the counts are still every launch
of a class, and the totals are the mean of the timed launches times that count."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


@pytest.mark.parametrize("every", [1, 4])
def test_profile_counts_every_launch_and_times_a_sample(fm, S, every):
    cfg = fm.configs.make(2, dims=(6, 5, 4))
    fm.configs.load(S, cfg)
    S.assemble(None, fm.FM_MIXED)
    S.use_tangent(True)
    b = np.random.default_rng(3).standard_normal(cfg["X"].size)
    b[cfg["dir_dofs"]] = 0.0
    S.profile(every)
    S.profile_reset()
    _, its, _ = S.pcg(b, tol=1e-30, max_it=10, prec=fm.FM_MIXED)
    prof = S.profile_get()
    S.profile(False)
    assert its == 10
    # the initial residual's product and one per iteration
    ms, n = prof["spmv"]
    assert n == 11 and ms > 0.0
    for k in ("ilu_fwd", "ilu_bwd"):
        ms, n = prof[k]
        assert n >= 10 and ms > 0.0, (k, prof[k])
    assert prof["ilu_factor"][1] >= 1
    # nothing is recorded while profiling is off
    S.profile_reset()
    S.pcg(b, tol=1e-30, max_it=3, prec=fm.FM_MIXED)
    assert all(n == 0 for _, n in S.profile_get().values())
