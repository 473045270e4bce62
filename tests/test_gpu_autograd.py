"""torch.autograd harness of the adjoint: gradients of a displacement loss
through the GPU Newton solve, checked against central finite differences
(step 1e-6, FP64) and against the FP64 CPU oracle's adjoint."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autograd_adjoint_matches_fd_and_oracle(O):
    torch = pytest.importorskip("torch")
    import flexmesh
    from flexmesh import autograd as FA

    cfg = flexmesh.configs.make(4, dims=(4, 3, 3))
    cfg["n_steps"] = 3
    nn = cfg["X"].size // 3
    top = cfg["loss_nodes"]
    ustar = np.random.default_rng(1).uniform(-0.02, 0.02, 3 * nn)
    prob = FA.Problem(cfg, restart=True)
    mu = torch.tensor(cfg["mu"], requires_grad=True)
    lam = torch.tensor(cfg["lam"], requires_grad=True)
    scale = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    w = torch.zeros(3 * nn, dtype=torch.float64)
    w.view(nn, 3)[torch.from_numpy(top)] = 1.0
    us = torch.from_numpy(ustar)

    def loss(m, l, s):
        u = FA.newton_solve(m, l, s, prob)
        return 0.5 * torch.sum(w * (u - us) ** 2)

    L = loss(mu, lam, scale)
    L.backward()
    rng = np.random.default_rng(2)
    h = 1e-6
    with torch.no_grad():
        for _ in range(4):
            v = torch.from_numpy(rng.standard_normal(mu.numel())) * mu
            q = torch.from_numpy(rng.standard_normal(lam.numel())) * lam
            fd = (loss(mu + h * v, lam + h * q, scale) - loss(mu - h * v, lam - h * q, scale)) / (2 * h)
            ad = torch.dot(mu.grad, v) + torch.dot(lam.grad, q)
            assert abs(float(fd - ad)) <= 1e-5 * abs(float(ad))
        fd = (loss(mu, lam, scale + h) - loss(mu, lam, scale - h)) / (2 * h)
        assert abs(float(fd - scale.grad)) <= 1e-5 * abs(float(scale.grad))
    # the oracle's FP64 adjoint at the oracle's own converged state
    P = O.NewtonProblem(cfg["X"], cfg["conn"], cfg["mu"], cfg["lam"], cfg["dir_dofs"], cfg["dir_vals"], None,
                        cfg["n_steps"], newton_rtol=1e-13, pcg_tol=1e-13, pcg_max=5000)
    uo = np.zeros(3 * nn)
    for s in range(1, cfg["n_steps"] + 1):
        P.step(s, uo)
    g = np.zeros(3 * nn)
    g.reshape(nn, 3)[top] = (uo - ustar).reshape(nn, 3)[top]
    dmo, dlo, dldo, _, _ = P.adjoint(uo, g)
    assert np.abs(mu.grad.numpy() - dmo).max() <= 1e-8 * np.abs(dmo).max()
    assert np.abs(lam.grad.numpy() - dlo).max() <= 1e-8 * np.abs(dlo).max()
    assert abs(float(scale.grad) - dldo) <= 1e-8 * abs(dldo)
