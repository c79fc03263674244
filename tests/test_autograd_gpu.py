"""Adjoint Greeks through torch.autograd: gradients equal the kernels' own
pathwise Greeks and agree with FP64 central finite differences on common
random numbers (same Philox stream)."""
import numpy as np
import pytest
import torch

import autograd as ag
import mcengine as mc

pytestmark = pytest.mark.gpu


def _fixed(n):
    one = np.ones(n)
    return dict(K=100 * one, T=one, kappa=2 * one, theta=.04 * one, xi=.3 * one, rho=-.7)


def test_heston_autograd_matches_fd(engine):
    n, paths, steps = 3, 1 << 15, 24
    m = mc.mode(mc.FP64, seed=21)
    fx = _fixed(n)
    S0 = torch.tensor([90.0, 100.0, 110.0], dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor([.04, .05, .03], dtype=torch.float64, requires_grad=True)
    r = torch.tensor(.03, dtype=torch.float64, requires_grad=True)
    price = ag.heston_asian_price(engine, S0, v0, r, fx, paths, steps, m)
    price.sum().backward()

    def p(S, v, rr):
        return engine.heston_asian(S, fx["K"], fx["T"], v, fx["kappa"], fx["theta"], fx["xi"], rr, fx["rho"], paths,
                                   steps, m)[0]

    S, v = S0.detach().numpy(), v0.detach().numpy()
    assert np.array_equal(price.detach().numpy(), p(S, v, .03))
    h = 1e-5
    fd_S = (p(S + h, v, .03) - p(S - h, v, .03)) / (2 * h)
    fd_v = (p(S, v + h * 1e-2, .03) - p(S, v - h * 1e-2, .03)) / (2 * h * 1e-2)
    fd_r = np.sum(p(S, v, .03 + h) - p(S, v, .03 - h)) / (2 * h)
    assert np.allclose(S0.grad.numpy(), fd_S, rtol=1e-3)
    assert np.allclose(v0.grad.numpy(), fd_v, rtol=1e-3)
    assert r.grad.item() == pytest.approx(fd_r, rel=1e-3)


def test_gbm_autograd_matches_fd(engine):
    n, paths = 4, 1 << 16
    m = mc.mode(mc.FP64, seed=22)
    K, T = np.array([20.0, 25.0, 30.0, 15.0]), np.array([1.0, 2.0, .5, 3.0])
    S0 = torch.tensor([22.0, 24.0, 28.0, 18.0], dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(.3, dtype=torch.float64, requires_grad=True)
    r = torch.tensor(.02, dtype=torch.float64, requires_grad=True)
    price = ag.gbm_european_price(engine, S0, sig, r, K, T, paths, m)
    (price * torch.tensor([1.0, 2.0, 3.0, 4.0], dtype=torch.float64)).sum().backward()
    w = np.array([1.0, 2.0, 3.0, 4.0])
    S = S0.detach().numpy()

    def p(S, s, rr):
        return engine.gbm_european(S, K, T, rr, s, paths, m)[0]

    h = 1e-6
    assert np.allclose(S0.grad.numpy(), w * (p(S + h, .3, .02) - p(S - h, .3, .02)) / (2 * h), rtol=1e-4)
    assert sig.grad.item() == pytest.approx(np.sum(w * (p(S, .3 + h, .02) - p(S, .3 - h, .02))) / (2 * h), rel=1e-4)
    assert r.grad.item() == pytest.approx(np.sum(w * (p(S, .3, .02 + h) - p(S, .3, .02 - h))) / (2 * h), rel=1e-4)


def test_libor_autograd_matches_fd(engine):
    paths = 1 << 14
    m = mc.mode(mc.FP64, seed=23)
    L0 = torch.tensor(.05, dtype=torch.float64, requires_grad=True)
    v, deltas = ag.libor_portfolio_value(engine, L0, .2, paths, m)
    v.backward()
    # the portfolio payoff has kinks (exercise boundaries): a 1e-5 bump moves
    # paths across them (FD off by 3e-5 relative), 1e-7 does not (1e-10)
    h = 1e-7
    fd = (engine.libor(paths, m, L0=.05 + h)[0] - engine.libor(paths, m, L0=.05 - h)[0]) / (2 * h)
    assert L0.grad.item() == pytest.approx(fd, rel=1e-7)
    assert L0.grad.item() == pytest.approx(deltas.sum().item(), rel=1e-15)
    # FP32 kernels through the same wrapper: within a small fraction of the FP64 value's SE
    v32, _ = ag.libor_portfolio_value(engine, torch.tensor(.05, dtype=torch.float64), .2, paths,
                                      mc.mode(mc.FP32, seed=23))
    se = engine.libor(paths, m)[1]
    assert abs(v32.item() - v.item()) < 0.05 * se
