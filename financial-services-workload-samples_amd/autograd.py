"""torch.autograd.Function wrappers over the engine's adjoint entry points.

The forward pass prices with the HIP kernels and keeps the pathwise Greeks
that the same launch produced (the Heston reverse sweep, the LIBOR adjoint
of path_calc_b2 / portfolio_b in liborSwaptionGreeks.cpp:86-211, the GBM
pathwise derivatives); the backward pass only scales them by the incoming
gradient.  Every Monte Carlo quantity is a mean over paths, so these are the
derivatives of the estimator itself, for a fixed Philox stream: a central
finite difference on the same seed converges to them.

    price = HestonAsianPrice.apply(engine, S0, v0, r, fixed, n_paths, n_steps, mode)
    price.sum().backward()      # S0.grad = delta, v0.grad = dP/dv0, r.grad = sum of rho
"""
import numpy as np
import torch

import mcengine as mc


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy()


class HestonAsianPrice(torch.autograd.Function):
    """Heston arithmetic-Asian call prices per option as a function of S0[n],
    v0[n] and the shared riskfree rate r (0-d).  `fixed` holds K, T, kappa,
    theta, xi (arrays of n) and rho."""

    @staticmethod
    def forward(ctx, engine, S0, v0, r, fixed, n_paths, n_steps, mode):
        res, _, g = engine.heston_asian(_np(S0), fixed["K"], fixed["T"], _np(v0), fixed["kappa"], fixed["theta"],
                                        fixed["xi"], float(r), fixed["rho"], n_paths, n_steps, mode, greeks=True)
        # vega is dP/dsqrt(v0); the chain rule gives dP/dv0 = vega / (2 sqrt(v0))
        dv0 = g[:, 1] / (2.0 * np.sqrt(_np(v0)))
        ctx.save_for_backward(torch.tensor(g[:, 0], dtype=S0.dtype, device=S0.device),
                              torch.tensor(dv0, dtype=v0.dtype, device=v0.device),
                              torch.tensor(g[:, 2], dtype=r.dtype, device=r.device))
        return torch.tensor(res, dtype=S0.dtype, device=S0.device)

    @staticmethod
    def backward(ctx, grad):
        dS, dv, dr = ctx.saved_tensors
        return None, grad * dS, grad * dv, (grad.to(dr.device) * dr).sum(), None, None, None, None


class GbmEuropeanPrice(torch.autograd.Function):
    """GBM European call prices per option as a function of S0[n], and of the
    shared volatility and riskfree rate (0-d tensors); K and T fixed."""

    @staticmethod
    def forward(ctx, engine, S0, sigma, r, K, T, n_paths, mode):
        res, _, g = engine.gbm_european(_np(S0), K, T, float(r), float(sigma), n_paths, mode, greeks=True)
        ctx.save_for_backward(torch.tensor(g[:, 0], dtype=S0.dtype, device=S0.device),
                              torch.tensor(g[:, 1], dtype=sigma.dtype, device=sigma.device),
                              torch.tensor(g[:, 2], dtype=r.dtype, device=r.device))
        return torch.tensor(res, dtype=S0.dtype, device=S0.device)

    @staticmethod
    def backward(ctx, grad):
        dS, dsig, dr = ctx.saved_tensors
        return None, grad * dS, (grad * dsig).sum(), (grad * dr).sum(), None, None, None, None


class LiborPortfolioValue(torch.autograd.Function):
    """Value of the 15-swaption portfolio (liborSwaptionGreeks.cpp) as a
    function of a flat initial forward curve L0 (0-d tensor): the gradient is
    the sum of the 80 adjoint deltas dV/dL0_i, which forward() also returns
    (without a graph) for callers that want the whole curve sensitivity."""

    @staticmethod
    def forward(ctx, engine, L0, lam, n_paths, mode):
        v, _, d, _ = engine.libor(n_paths, mode, L0=float(L0), lam=lam)
        deltas = torch.tensor(d, dtype=L0.dtype, device=L0.device)
        ctx.save_for_backward(deltas.sum())
        ctx.mark_non_differentiable(deltas)
        return torch.tensor(v, dtype=L0.dtype, device=L0.device), deltas

    @staticmethod
    def backward(ctx, grad, _grad_deltas):
        (dsum,) = ctx.saved_tensors
        return None, grad * dsum, None, None, None


def heston_asian_price(engine, S0, v0, r, fixed, n_paths, n_steps, mode=None):
    return HestonAsianPrice.apply(engine, S0, v0, r, fixed, n_paths, n_steps, mode or mc.mode(mc.FP32))


def gbm_european_price(engine, S0, sigma, r, K, T, n_paths, mode=None):
    return GbmEuropeanPrice.apply(engine, S0, sigma, r, K, T, n_paths, mode or mc.mode(mc.FP32))


def libor_portfolio_value(engine, L0, lam=0.2, n_paths=262144, mode=None):
    return LiborPortfolioValue.apply(engine, L0, lam, n_paths, mode or mc.mode(mc.FP32))
