"""GBM European on the GPU against the CPU oracle and Black-Scholes.

Tolerances: FP64 buffer/Philox modes agree with the oracle to <= 1e-12
relative (same inputs, same arithmetic up to summation order); FP32 modes to
<= 1e-5 relative; statistical checks use the estimator's own standard error.
"""
import math

import numpy as np
import pytest

import mcengine as mc
import oracle as orc

pytestmark = pytest.mark.gpu

L2E = 1.4426950408889634074


def _opts(n, seed=0):
    rng = np.random.default_rng(seed)
    # config-0 distributions: S0~U[5,30], K~U[1,100], T~U[0.25,10]
    return rng.uniform(5, 30, n), rng.uniform(1, 100, n), rng.uniform(.25, 10, n)


def _itm_opts(n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(20, 30, n), rng.uniform(5, 20, n), rng.uniform(.5, 3, n)


def _samples(S, X, T, r, sig, n_paths, seed):
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((len(S), n_paths))
    mu = (L2E * (r - .5 * sig * sig) * T)[:, None]
    vol = (L2E * sig * np.sqrt(T))[:, None]
    return (mu + vol * z).astype(np.float32)


def _rel(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def test_buffer_fp64_matches_oracle(engine):
    S, X, T = _opts(12, 1)
    r, sig, n_paths = 0.02, 0.3, 8192
    sm = _samples(S, X, T, r, sig, n_paths, 2)
    res, conf = engine.gbm_european(S, X, T, r, sig, n_paths, mc.mode(mc.FP64, mc.RNG_BUFFER), samples=sm)
    ores, oconf = orc.gbm_buffer(S, X, T, r, sig, sm)
    nz = ores > 0
    assert _rel(res[nz], ores[nz]) <= 1e-12
    assert _rel(conf[nz], oconf[nz]) <= 1e-12
    assert np.all(res[~nz] == 0)


def test_buffer_fp32_matches_oracle(engine):
    S, X, T = _itm_opts(8, 3)
    r, sig, n_paths = 0.02, 0.3, 16384
    sm = _samples(S, X, T, r, sig, n_paths, 4)
    res, conf = engine.gbm_european(S, X, T, r, sig, n_paths, mc.mode(mc.FP32, mc.RNG_BUFFER), samples=sm)
    ores, oconf = orc.gbm_buffer(S, X, T, r, sig, sm)
    assert _rel(res, ores) <= 1e-5
    assert _rel(conf, oconf) <= 1e-4


@pytest.mark.parametrize("n_paths", [4096, 1001, 2])
def test_philox_fp64_matches_oracle(engine, n_paths):
    S, X, T = _itm_opts(6, 5)
    r, sig = 0.02, 0.3
    res, conf, g = engine.gbm_european(S, X, T, r, sig, n_paths, mc.mode(mc.FP64, seed=77, stream=3),
                                       greeks=True)
    ores, oconf, og = orc.gbm_philox(S, X, T, r, sig, n_paths, seed=77, stream=3, greeks=True)
    assert _rel(res, ores) <= 1e-12
    if n_paths > 2:
        assert _rel(conf, oconf) <= 1e-11
        assert _rel(g[:, :3], og[:, :3]) <= 1e-12
        assert _rel(g[:, 3:], og[:, 3:]) <= 1e-10


def test_philox_fp32_vs_fp64(engine):
    S, X, T = _itm_opts(16, 6)
    r, sig, n = 0.02, 0.3, 1 << 16
    r32, c32, g32 = engine.gbm_european(S, X, T, r, sig, n, mc.mode(mc.FP32, seed=9), greeks=True)
    r64, c64, g64 = engine.gbm_european(S, X, T, r, sig, n, mc.mode(mc.FP64, seed=9), greeks=True)
    # same Philox stream: only FP32 arithmetic (and hardware transcendentals) differ
    assert _rel(r32, r64) <= 1e-5
    # Greeks: relative 1e-5, or 1e-3 of their standard error where a Greek is
    # near zero (vega of deep in-the-money calls cancels)
    err = np.abs(g32[:, :3] - g64[:, :3])
    assert np.all(err <= 1e-5 * np.abs(g64[:, :3]) + 1e-3 * g64[:, 3:])


def test_1024x262144_against_black_scholes(engine):
    # Full shape: 1024 options x 262144 paths, FP32 Philox.
    S, X, T = _opts(1024, 10)
    r, sig, n = 0.02, 0.3, 262144
    res, conf, g = engine.gbm_european(S, X, T, r, sig, n, mc.mode(mc.FP32, seed=2024), greeks=True)
    se = conf / 1.96
    bs = np.array([orc.bs_call(S[i], X[i], T[i], r, sig) for i in range(len(S))])
    # the z-score is only Gaussian with enough in-the-money paths
    live = (res > 0) & ((res / np.where(se > 0, se, 1.0)) ** 2 > 50)
    z = (res[live] - bs[live]) / se[live]
    assert live.sum() > 700
    assert np.mean(np.abs(z) > 3) < 0.01  # ~0.27 % expected
    assert np.max(np.abs(z)) < 5.5
    assert abs(np.mean(z * z) - 1.0) < 0.2
    # the rest are far out of the money
    assert np.all(np.abs(res[~live] - bs[~live]) < 1e-2)
    gb = np.array([orc.bs_call_greeks(S[i], X[i], T[i], r, sig) for i in range(len(S))])
    for k in range(3):
        ok = live & (g[:, 3 + k] > 0)
        zk = (g[ok, k] - gb[ok, k]) / g[ok, 3 + k]
        assert np.mean(np.abs(zk) > 3) < 0.01 and np.max(np.abs(zk)) < 5.5


def test_crr_lattice_agreement(engine):
    S, X, T = _itm_opts(4, 12)
    r, sig = 0.02, 0.3
    res, conf = engine.gbm_european(S, X, T, r, sig, 1 << 20, mc.mode(mc.FP64, seed=5))
    for i in range(4):
        crr = orc.crr_european_call(S[i], X[i], T[i], r, sig, 2048)
        assert abs(res[i] - crr) < 4 * conf[i] / 1.96 + 2e-3 * crr


def test_deterministic(engine):
    S, X, T = _opts(32, 13)
    a = engine.gbm_european(S, X, T, .02, .3, 50000, mc.mode(mc.FP32, seed=1))
    b = engine.gbm_european(S, X, T, .02, .3, 50000, mc.mode(mc.FP32, seed=1))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    c = engine.gbm_european(S, X, T, .02, .3, 50000, mc.mode(mc.FP32, seed=1, stream=1))
    assert not np.array_equal(a[0], c[0])


def test_async_matches_sync(engine):
    import torch
    S, X, T = _opts(64, 14)
    m = mc.mode(mc.FP32, seed=4)
    res, conf, g = engine.gbm_european(S, X, T, .02, .3, 100000, m, greeks=True)
    dev = torch.device("cuda:0")
    tS, tX, tT = (torch.tensor(a, dtype=torch.float64, device=dev) for a in (S, X, T))
    out = torch.zeros(64, dtype=torch.float64, device=dev)
    oc = torch.zeros_like(out)
    og = torch.zeros(64, 6, dtype=torch.float64, device=dev)
    ws = torch.empty(mc.gbm_european_workspace(64, 100000, .02, .3, m), dtype=torch.uint8, device=dev)
    mc.gbm_european_async(tS, tX, tT, .02, .3, 100000, m, out, oc, ws, out_greeks=og)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), res)
    assert np.array_equal(oc.cpu().numpy(), conf)
    assert np.array_equal(og.cpu().numpy(), g)


def test_errors(engine):
    S, X, T = _opts(4, 15)
    with pytest.raises(mc.FsmcError):
        engine.gbm_european(S, X, T, .02, -.3, 1000, mc.mode())
    with pytest.raises(mc.FsmcError):
        engine.gbm_european(S, X, T, .02, .3, 1, mc.mode())
    with pytest.raises(mc.FsmcError):  # buffer mode without samples
        engine.gbm_european(S, X, T, .02, .3, 1000, mc.mode(rng=mc.RNG_BUFFER))


# ---- control variate (S_T; an engine option, no counterpart in the sample) ----
def test_control_variate_fp64_matches_oracle(engine):
    S, X, T = _opts(16, 7)
    r, sig, n_paths = 0.02, 0.3, 10000  # not a multiple of 4: a partial Philox group
    res, conf = engine.gbm_european(S, X, T, r, sig, n_paths, mc.mode(mc.FP64, seed=8), control_variate=True)
    ores, oconf = orc.gbm_philox_cv(S, X, T, r, sig, n_paths, seed=8)
    _, pconf = orc.gbm_philox(S, X, T, r, sig, n_paths, seed=8)
    nz = ores > 0
    assert _rel(res[nz], ores[nz]) <= 1e-12 and np.all(res[~nz] == 0)
    # var(y) - beta cov cancels deep in the money, and the confidence is its
    # square root: a summation-order difference of 1e-16 var(y) becomes
    # 1e-8 of the plain estimator's confidence
    assert np.all(np.abs(conf - oconf) <= 2e-7 * pconf + 1e-13 * S)
    sm = _samples(S, X, T, r, sig, 4096, 9)
    res, conf = engine.gbm_european(S, X, T, r, sig, 4096, mc.mode(mc.FP64, mc.RNG_BUFFER), samples=sm,
                                    control_variate=True)
    ores, oconf = orc.gbm_buffer_cv(S, X, T, r, sm)
    _, pconf = orc.gbm_buffer(S, X, T, r, sig, sm)
    nz = ores > 0
    assert _rel(res[nz], ores[nz]) <= 1e-12 and np.all(res[~nz] == 0)
    assert np.all(np.abs(conf - oconf) <= 2e-7 * pconf + 1e-13 * S)


def test_control_variate_small_strike_exact_and_greeks_unchanged(engine):
    S, X, T = np.array([20.0, 7.5, 12.0]), np.array([1e-3, 2e-3, 5e-3]), np.array([1.0, 3.0, .5])
    r, sig = 0.02, 0.3
    exact = S - X * np.exp(-r * T)
    res, conf, g = engine.gbm_european(S, X, T, r, sig, 65536, mc.mode(mc.FP64, seed=2), greeks=True,
                                       control_variate=True)
    assert np.allclose(res, exact, rtol=1e-12, atol=0) and np.all(conf < 1e-9 * S)
    _, _, g0 = engine.gbm_european(S, X, T, r, sig, 65536, mc.mode(mc.FP64, seed=2), greeks=True)
    assert np.array_equal(g, g0)
    res32, conf32 = engine.gbm_european(S, X, T, r, sig, 65536, mc.mode(mc.FP32, seed=2), control_variate=True)
    assert np.allclose(res32, exact, rtol=1e-5, atol=0)


def test_control_variate_fp32_and_variance_reduction(engine):
    # the 1024 x 262144 workload: FP32 CV vs FP64 CV on the same stream, CV vs
    # Black-Scholes, and the standard error against the plain estimator's
    rng = np.random.default_rng(11)
    n = 1024
    S, X, T = rng.uniform(5, 30, n), rng.uniform(1, 100, n), rng.uniform(.25, 10, n)
    r, sig, n_paths = 0.02, 0.3, 262144
    c64, f64 = engine.gbm_european(S, X, T, r, sig, n_paths, mc.mode(mc.FP64, seed=4), control_variate=True)
    c32, f32 = engine.gbm_european(S, X, T, r, sig, n_paths, mc.mode(mc.FP32, seed=4), control_variate=True)
    p64, g64 = engine.gbm_european(S, X, T, r, sig, n_paths, mc.mode(mc.FP64, seed=4))
    se, pse = f64 / 1.96, g64 / 1.96
    bs = np.array([orc.bs_call(*a, r, sig) for a in zip(S, X, T)])
    # the variance of the CV estimator comes from the paths on both sides of
    # the strike; its standard error is a fair yardstick only with enough of
    # each (expected counts from P(S_T > K) = N(d2))
    d2 = (np.log(S / X) + (r - .5 * sig * sig) * T) / (sig * np.sqrt(T))
    p_itm = np.array([.5 * math.erfc(-x / math.sqrt(2)) for x in d2])
    both = (n_paths * p_itm >= 1000) & (n_paths * (1 - p_itm) >= 1000)
    z = (c64[both] - bs[both]) / se[both]
    assert both.sum() > 300 and np.max(np.abs(z)) < 5.0 and abs(np.mean(z * z) - 1.0) < 0.2
    # far from the money one side may see no path at all (either estimator
    # then misses the call value, or the put value deep in the money)
    put = bs - (S - X * np.exp(-r * T))
    slack = np.where(p_itm < .5, bs, np.maximum(put, 0.0))
    assert np.all(np.abs(c64[~both] - bs[~both]) <= 5 * pse[~both] + slack[~both] + 1e-9 * S[~both])
    assert np.all(np.abs(c32 - c64) <= 0.05 * se + 2e-6 * S)
    assert np.all(f64 <= g64 * (1 + 1e-9))
    itm = S > 1.2 * X
    assert np.median(f64[itm] / g64[itm]) < 0.3
