"""Heston arithmetic-Asian on the GPU against the CPU oracle.

Buffer (oracle) mode consumes normals in the reference's consumption layout
[option][chunk of 64][step][lane] and must match the oracle's restatement of
MonteCarloTimeStep to <= 1e-12 relative in FP64 (<= 1e-5 in FP32), with and
without each reference quirk.  Philox mode must match the oracle's Philox
run in FP64; the adjoint Greeks match the oracle's tangent-mode Greeks and
central finite differences.
"""
import math

import numpy as np
import pytest

import mcengine as mc
import oracle as orc

pytestmark = pytest.mark.gpu

Q4, Q5 = mc.QUIRK_ASIAN_VAR_AS_VOL, mc.QUIRK_ASIAN_RUNNING_VAL2


def _params(n, seed=0, ref_ranges=False):
    rng = np.random.default_rng(seed)
    if ref_ranges:  # MCAsianOptions.cpp:255-266 parameter ranges
        return dict(S=rng.uniform(5, 50, n), X=rng.uniform(10, 25, n), T=rng.uniform(1, 5, n),
                    v0=rng.uniform(.01, .015, n), kappa=rng.uniform(6, 7, n), theta=rng.uniform(.02, .03, n),
                    xi=rng.uniform(.45, .6, n))
    return dict(S=rng.uniform(80, 120, n), X=np.full(n, 100.0), T=np.ones(n), v0=np.full(n, .04),
                kappa=np.full(n, 2.0), theta=np.full(n, .04), xi=np.full(n, .3))


def _rel(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def _call(engine, P, r, rho, n_paths, n_steps, m, **kw):
    return engine.heston_asian(P["S"], P["X"], P["T"], P["v0"], P["kappa"], P["theta"], P["xi"], r, rho,
                               n_paths, n_steps, m, **kw)


def _orc_buf(P, r, rho, n_paths, n_steps, z1, z2, q):
    return orc.heston_buffer(P["S"], P["X"], P["T"], P["v0"], P["kappa"], P["theta"], P["xi"], r, rho, n_paths,
                             n_steps, z1, z2, quirks=q)


@pytest.mark.parametrize("quirks", [0, Q4, Q5, Q4 | Q5])
def test_buffer_fp64_matches_oracle(engine, quirks):
    n_opt, n_paths, n_steps = 5, 256, 128
    P = _params(n_opt, 1, ref_ranges=True)
    rng = np.random.default_rng(2)
    z1 = rng.standard_normal(n_opt * n_paths * n_steps)
    z2 = rng.standard_normal(n_opt * n_paths * n_steps)
    r, rho = 0.03, -0.65
    res, conf = _call(engine, P, r, rho, n_paths, n_steps, mc.mode(mc.FP64, mc.RNG_BUFFER, quirks=quirks), z1=z1,
                      z2=z2)
    oq = (orc.Q_VAR_AS_VOL if quirks & Q4 else 0) | (orc.Q_RUNNING_VAL2 if quirks & Q5 else 0)
    ores, oconf = _orc_buf(P, r, rho, n_paths, n_steps, z1, z2, oq)
    assert _rel(res, ores) <= 1e-12
    assert _rel(conf, oconf) <= 1e-12


def test_buffer_fp32_matches_oracle(engine):
    n_opt, n_paths, n_steps = 4, 512, 64
    P = _params(n_opt, 3)
    rng = np.random.default_rng(4)
    z1 = rng.standard_normal(n_opt * n_paths * n_steps)
    z2 = rng.standard_normal(n_opt * n_paths * n_steps)
    res, conf = _call(engine, P, .03, -.7, n_paths, n_steps, mc.mode(mc.FP32, mc.RNG_BUFFER), z1=z1, z2=z2)
    ores, oconf = _orc_buf(P, .03, -.7, n_paths, n_steps, z1, z2, 0)
    assert _rel(res, ores) <= 1e-5
    assert _rel(conf, oconf) <= 1e-4


@pytest.mark.parametrize("n_paths,n_steps", [(1000, 7), (256, 2), (64, 1)])
def test_philox_fp64_matches_oracle(engine, n_paths, n_steps):
    P = _params(3, 5)
    res, conf = _call(engine, P, .03, -.7, n_paths, n_steps, mc.mode(mc.FP64, seed=31, stream=2))
    ores, oconf = orc.heston_philox(P["S"], P["X"], P["T"], P["v0"], P["kappa"], P["theta"], P["xi"], .03, -.7,
                                    n_paths, n_steps, seed=31, stream=2)
    assert _rel(res, ores) <= 1e-12
    assert _rel(conf, oconf) <= 1e-10


def test_philox_fp32_vs_fp64_64x131072x365(engine):
    # Full shape: 64 options x 131072 paths x 365 steps.
    P = _params(64, 6)
    r32, c32 = _call(engine, P, .03, -.7, 131072, 365, mc.mode(mc.FP32, seed=1))
    r64, c64 = _call(engine, P, .03, -.7, 131072, 365, mc.mode(mc.FP64, seed=1))
    # identical Philox stream: the FP32 error (measured ~3e-7 of spot after
    # 365 steps) is bounded by 1e-6 of spot and is far below one standard error
    err = np.abs(r32 - r64)
    assert np.all(err <= 1e-6 * P["S"])
    assert np.all(err < 0.01 * c64 / 1.96)
    # K = 0 pays the average: exact expectation e^{-rT} S0 mean_k e^{r k dt}
    P0 = dict(P, X=np.zeros(64))
    r0, c0 = _call(engine, P0, .03, -.7, 131072, 365, mc.mode(mc.FP32, seed=2))
    dt = 1 / 365
    exact = np.exp(-.03) * P["S"] * np.mean(np.exp(.03 * dt * np.arange(1, 366)))
    z = (r0 - exact) / (c0 / 1.96)
    assert np.max(np.abs(z)) < 5 and abs(np.mean(z * z) - 1) < 0.5


@pytest.mark.parametrize("precision,var_as_vol", [(mc.FP64, False), (mc.FP32, False), (mc.FP32, True)])
def test_two_step_quadrature(engine, precision, var_as_vol):
    from test_oracle_pins import _heston_n2_quadrature
    P = dict(S=[100.0], X=[100.0], T=[1.0], v0=[.09], kappa=[2.0], theta=[.04], xi=[1.0])
    q = Q4 if var_as_vol else 0
    res, conf = _call(engine, P, .03, -.7, 1 << 22, 2, mc.mode(precision, seed=8, quirks=q))
    exact = _heston_n2_quadrature(100, 100, 1, .03, .09, 2, .04, 1.0, -.7, var_as_vol)
    assert abs(res[0] - exact) < 4 * conf[0] / 1.96


def test_greeks_fp64_match_oracle_tangents(engine):
    P = _params(4, 7)
    n_paths, n_steps = 2000, 37  # partial last checkpoint segment, odd steps
    res, conf, g = _call(engine, P, .03, -.7, n_paths, n_steps, mc.mode(mc.FP64, seed=12), greeks=True)
    ores, oconf, og = orc.heston_philox(P["S"], P["X"], P["T"], P["v0"], P["kappa"], P["theta"], P["xi"], .03,
                                        -.7, n_paths, n_steps, seed=12, greeks=True)
    assert _rel(res, ores) <= 1e-12
    assert _rel(g[:, :3], og[:, :3]) <= 1e-9
    assert _rel(g[:, 3:], og[:, 3:]) <= 1e-7
    # the adjoint kernel's forward sweep prices exactly like the forward kernel
    fres, fconf = _call(engine, P, .03, -.7, n_paths, n_steps, mc.mode(mc.FP64, seed=12))
    assert np.array_equal(fres, res)


def test_greeks_vs_finite_differences(engine):
    # FP64 adjoint vs FP64 central differences (common random numbers).
    P = _params(1, 8)
    P["S"] = np.array([100.0])
    n_paths, n_steps, r = 1 << 16, 52, .03
    m = mc.mode(mc.FP64, seed=13)
    res, conf, g = _call(engine, P, r, -.7, n_paths, n_steps, m, greeks=True)

    def price(dS=0.0, dsig=0.0, dr=0.0):
        Q = dict(P, S=P["S"] + dS, v0=(np.sqrt(P["v0"]) + dsig) ** 2)
        return _call(engine, Q, r + dr, -.7, n_paths, n_steps, m)[0][0]

    h = 1e-5
    fd = [(price(dS=h) - price(dS=-h)) / (2 * h), (price(dsig=h) - price(dsig=-h)) / (2 * h),
          (price(dr=h) - price(dr=-h)) / (2 * h)]
    assert np.allclose(g[0, :3], fd, rtol=1e-3)


def test_greeks_fp32_vs_fp64(engine):
    P = _params(8, 9)
    _, _, g32 = _call(engine, P, .03, -.7, 1 << 15, 365, mc.mode(mc.FP32, seed=3), greeks=True)
    _, _, g64 = _call(engine, P, .03, -.7, 1 << 15, 365, mc.mode(mc.FP64, seed=3), greeks=True)
    # FP32 Greek error far below their standard error
    assert np.all(np.abs(g32[:, :3] - g64[:, :3]) < 0.02 * g64[:, 3:])


def test_deterministic_and_async(engine):
    import torch
    P = _params(8, 10)
    m = mc.mode(mc.FP32, seed=5)
    a = _call(engine, P, .03, -.7, 4096, 50, m, greeks=True)
    b = _call(engine, P, .03, -.7, 4096, 50, m, greeks=True)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    dev = torch.device("cuda:0")
    prm = [torch.tensor(P[k], dtype=torch.float64, device=dev) for k in ("S", "X", "T", "v0", "kappa", "theta",
                                                                          "xi")]
    out = torch.zeros(8, dtype=torch.float64, device=dev)
    oc = torch.zeros_like(out)
    og = torch.zeros(8, 6, dtype=torch.float64, device=dev)
    ws = torch.empty(mc.heston_asian_workspace(8, 4096, 50, .03, -.7, m), dtype=torch.uint8, device=dev)
    mc.heston_asian_async(prm, 4096, 50, .03, -.7, m, out, oc, ws, out_greeks=og)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), a[0]) and np.array_equal(og.cpu().numpy(), a[2])


def test_errors(engine):
    P = _params(2, 11)
    with pytest.raises(mc.FsmcError):  # Greeks need Philox without quirks
        _call(engine, P, .03, -.7, 128, 10, mc.mode(quirks=Q4), greeks=True)
    with pytest.raises(mc.FsmcError):
        _call(engine, P, .03, -.7, 128, 0, mc.mode())


# ---- forward (tangent) Greeks: fsmc_heston_asian_tangent_async ----
def _tangent(P, r, rho, n_paths, n_steps, m):
    import torch
    dev = torch.device("cuda:0")
    prm = [torch.tensor(P[k], dtype=torch.float64, device=dev) for k in ("S", "X", "T", "v0", "kappa", "theta",
                                                                          "xi")]
    n = len(P["S"])
    out = torch.zeros(n, dtype=torch.float64, device=dev)
    oc = torch.zeros_like(out)
    og = torch.zeros(n, 6, dtype=torch.float64, device=dev)
    ws = torch.empty(mc.heston_asian_tangent_workspace(n, n_paths, n_steps, r, rho, m), dtype=torch.uint8,
                     device=dev)
    mc.heston_asian_tangent_async(prm, n_paths, n_steps, r, rho, m, out, oc, og, ws)
    torch.cuda.synchronize()
    return out.cpu().numpy(), oc.cpu().numpy(), og.cpu().numpy()


def test_tangent_fp64_matches_oracle_and_adjoint(engine):
    # the oracle computes these Greeks by the same forward recursion: FP64
    # agrees to rounding; the adjoint (reverse sweep) to its own rounding
    P = _params(4, 7)
    n_paths, n_steps = 2000, 37
    m = mc.mode(mc.FP64, seed=12)
    res, conf, g = _tangent(P, .03, -.7, n_paths, n_steps, m)
    ores, oconf, og = orc.heston_philox(P["S"], P["X"], P["T"], P["v0"], P["kappa"], P["theta"], P["xi"], .03,
                                        -.7, n_paths, n_steps, seed=12, greeks=True)
    assert _rel(res, ores) <= 1e-12
    assert _rel(g[:, :3], og[:, :3]) <= 1e-11
    assert _rel(g[:, 3:], og[:, 3:]) <= 1e-9
    _, _, ga = _call(engine, P, .03, -.7, n_paths, n_steps, m, greeks=True)
    assert _rel(g[:, :3], ga[:, :3]) <= 1e-9
    # the path is stepped exactly as the forward kernel steps it (the
    # confidence goes through the Greeks finalize's arithmetic: to rounding)
    fres, fconf = _call(engine, P, .03, -.7, n_paths, n_steps, m)
    assert np.array_equal(fres, res) and _rel(conf, fconf) <= 1e-13


def test_tangent_fp32(engine):
    P = _params(8, 9)
    n_paths, n_steps = 1 << 15, 365
    r32, c32, g32 = _tangent(P, .03, -.7, n_paths, n_steps, mc.mode(mc.FP32, seed=3))
    _, _, g64 = _tangent(P, .03, -.7, n_paths, n_steps, mc.mode(mc.FP64, seed=3))
    # FP32 Greek error far below their standard error, and the FP32 adjoint's
    assert np.all(np.abs(g32[:, :3] - g64[:, :3]) < 0.02 * g64[:, 3:])
    _, _, ga32 = _call(engine, P, .03, -.7, n_paths, n_steps, mc.mode(mc.FP32, seed=3), greeks=True)
    assert np.all(np.abs(g32[:, :3] - ga32[:, :3]) < 0.02 * g64[:, 3:])
    # FP32 price: bit for bit the forward kernel's
    f32, fc32 = _call(engine, P, .03, -.7, n_paths, n_steps, mc.mode(mc.FP32, seed=3))
    assert np.array_equal(f32, r32) and _rel(c32, fc32) <= 1e-13
    # deterministic
    assert all(np.array_equal(a, b) for a, b in zip(_tangent(P, .03, -.7, n_paths, n_steps,
                                                              mc.mode(mc.FP32, seed=3)), (r32, c32, g32)))


def test_tangent_errors(engine):
    import torch
    P = _params(2, 11)
    dev = torch.device("cuda:0")
    prm = [torch.tensor(P[k], dtype=torch.float64, device=dev) for k in ("S", "X", "T", "v0", "kappa", "theta",
                                                                          "xi")]
    out, oc = torch.zeros(2, dtype=torch.float64, device=dev), torch.zeros(2, dtype=torch.float64, device=dev)
    og = torch.zeros(2, 6, dtype=torch.float64, device=dev)
    ws = torch.empty(1 << 16, dtype=torch.uint8, device=dev)
    with pytest.raises(mc.FsmcError):  # Greeks need Philox without quirks
        mc.heston_asian_tangent_async(prm, 128, 10, .03, -.7, mc.mode(quirks=Q4), out, oc, og, ws)
    with pytest.raises(mc.FsmcError):
        mc.heston_asian_tangent_async(prm, 128, 10, .03, -.7, mc.mode(), out, oc, None, ws)
    with pytest.raises(mc.FsmcError):  # workspace too small
        mc.heston_asian_tangent_async(prm, 1 << 20, 10, .03, -.7, mc.mode(), out, oc, og, ws[:64])
