"""Longstaff-Schwartz on the GPU.

Reference algorithm (MonteCarloAmericanOptions) in buffer mode matches the
oracle's restatement to <= 1e-12 relative in FP64.  The batched LSMC matches
the oracle's FP64 LSMC to <= 1e-12 on the same Philox stream; its FP32
build is checked against the FP64 GPU run and against a Bermudan CRR lattice
with the same exercise dates (LSMC is biased low by its regression)."""
import math

import numpy as np
import pytest

import mcengine as mc
import oracle as orc

pytestmark = pytest.mark.gpu


def _philox_z(n_paths, n_steps, seed, option=0):
    # engine contract, LSMC tag 4: counter (path, step/4, option)
    z = np.empty((n_steps, n_paths))
    for p in range(n_paths):
        for q in range((n_steps + 3) // 4):
            n = orc.normals4(seed, 0, 4, p, q, option)
            for j in range(4):
                if 4 * q + j < n_steps:
                    z[4 * q + j, p] = n[j]
    return z


@pytest.mark.parametrize("n_paths,n_steps", [(8192, 64), (4096, 256), (1000, 9)])
def test_reference_algorithm_fp64_buffer(engine, n_paths, n_steps):
    z = np.random.default_rng(n_steps).standard_normal(n_paths * n_steps)
    am, eu = engine.american_lsmc_ref(n_paths, n_steps, 100.0, 110.0, .25, 2.0, mc.mode(mc.FP64, mc.RNG_BUFFER), z)
    oam, oeu = orc.lsmc_ref(n_paths, n_steps, 100.0, 110.0, .25, 2.0, z)
    assert am == pytest.approx(oam, rel=1e-12)
    assert eu == pytest.approx(oeu, rel=1e-12)


def test_reference_algorithm_fp64_philox(engine):
    n_paths, n_steps = 2048, 32
    am, eu = engine.american_lsmc_ref(n_paths, n_steps, 100.0, 110.0, .25, 2.0, mc.mode(mc.FP64, seed=9))
    z = _philox_z(n_paths, n_steps, 9)
    oam, oeu = orc.lsmc_ref(n_paths, n_steps, 100.0, 110.0, .25, 2.0, z)
    assert am == pytest.approx(oam, rel=1e-12) and eu == pytest.approx(oeu, rel=1e-12)


def test_reference_algorithm_fp32(engine):
    n_paths, n_steps = 16384, 256
    z = np.random.default_rng(1).standard_normal(n_paths * n_steps)
    am, eu = engine.american_lsmc_ref(n_paths, n_steps, 100.0, 110.0, .25, 2.0, mc.mode(mc.FP32, mc.RNG_BUFFER), z)
    oam, oeu = orc.lsmc_ref(n_paths, n_steps, 100.0, 110.0, .25, 2.0, z)
    # FP32 paths move a handful of exercise decisions at the regression boundary
    assert eu == pytest.approx(oeu, rel=1e-5)
    assert am == pytest.approx(oam, rel=2e-3)


def _opts(n, seed):
    rng = np.random.default_rng(seed)
    return dict(S0=rng.uniform(80, 120, n), K=np.full(n, 100.0), sigma=rng.uniform(.1, .5, n), r=np.full(n, .03),
                T=rng.uniform(.25, 2, n), steps=rng.integers(50, 251, n), paths=rng.integers(16384, 262145, n))


def _call(engine, P, m):
    return engine.american_lsmc(P["S0"], P["K"], P["sigma"], P["r"], P["T"], P["steps"], P["paths"], m)


def test_batched_fp64_matches_oracle(engine):
    P = dict(S0=[90.0, 100.0, 110.0, 95.0], K=[100.0] * 4, sigma=[.2, .3, .4, .15], r=[.03] * 4,
             T=[1.0, .5, 2.0, .25], steps=[20, 9, 13, 2], paths=[3000, 1025, 2048, 700])
    p, se, eu = _call(engine, P, mc.mode(mc.FP64, seed=21))
    op, ose, oeu = orc.lsmc(P["S0"], P["K"], P["sigma"], P["r"], P["T"], P["steps"], P["paths"], seed=21)
    assert np.allclose(p, op, rtol=1e-12, atol=0)
    assert np.allclose(se, ose, rtol=1e-9, atol=0)
    assert np.allclose(eu, oeu, rtol=1e-12, atol=0)


def test_batched_longstaff_schwartz_table(engine):
    # Longstaff & Schwartz (2001) Table 1, LSM column, 50 dates, 100k paths
    P = dict(S0=[36.0, 40.0, 44.0], K=[40.0] * 3, sigma=[.2] * 3, r=[.06] * 3, T=[1.0] * 3, steps=[50] * 3,
             paths=[100000] * 3)
    for prec in (mc.FP32, mc.FP64):
        p, se, eu = _call(engine, P, mc.mode(prec, seed=5))
        assert p == pytest.approx([4.472, 2.313, 1.110], abs=0.03)


def test_256_uneven_puts_against_bermudan_lattice(engine):
    # 256 puts, steps ~U{50..250}, paths ~U{16384..262144}, FP32.
    P = _opts(256, 3)
    p, se, eu = _call(engine, P, mc.mode(mc.FP32, seed=3))
    p64, se64, eu64 = _call(engine, P, mc.mode(mc.FP64, seed=3))
    assert np.all(np.isfinite(p)) and np.all(se > 0)
    # FP32 paths (FP64 moment sums) against FP64 on the same stream: rounding
    # moves a few exercise decisions at the regression boundary; measured
    # median 0.05 SE, worst of 256 options 0.71 SE
    d = np.abs(p - p64) / se64
    assert np.median(d) < 0.15 and np.max(d) < 1.5
    lows = []
    for i in range(0, 256, 8):  # a 32-option sample of the lattice (4096+ steps each)
        n = int(P["steps"][i])
        berm = orc.crr_bermudan_put(P["S0"][i], 100.0, P["T"][i], .03, P["sigma"][i], n, max(1, 4096 // n + 1))
        # LSMC is a lower bound up to noise, and its regression bias is small
        assert p[i] < berm + 3 * se[i] + 1e-3
        lows.append((berm - p[i]) / berm)
        eu_bs = orc.bs_call(P["S0"][i], 100, P["T"][i], .03, P["sigma"][i]) - P["S0"][i] + 100 * math.exp(
            -.03 * P["T"][i])
        assert abs(eu[i] - eu_bs) < 4 * se[i] + 2e-3 * eu_bs
    assert np.median(lows) < 0.01 and max(lows) < 0.03


def test_plan_reruns_are_identical(engine):
    P = _opts(12, 4)
    P["paths"] = np.minimum(P["paths"], 40000)
    m = mc.mode(mc.FP32, seed=7)
    plan = engine.lsmc_plan(P["S0"], P["K"], P["sigma"], P["r"], P["T"], P["steps"], P["paths"], m)
    plan.run()
    a = plan.results()
    plan.run()
    b = plan.results()
    plan.close()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    c = _call(engine, P, m)
    assert all(np.array_equal(x, y) for x, y in zip(a, c))


def test_edge_cases(engine):
    # one step (European with exercise at 0), deep in/out of the money
    P = dict(S0=[100.0, 50.0, 200.0], K=[100.0] * 3, sigma=[.2] * 3, r=[.03] * 3, T=[1.0] * 3, steps=[1, 30, 30],
             paths=[20000, 5000, 5000])
    p, se, eu = _call(engine, P, mc.mode(mc.FP64, seed=2))
    assert p[0] == pytest.approx(max(0.0, eu[0]), rel=1e-12)
    assert p[1] == pytest.approx(50.0, abs=1e-9)  # immediate exercise is optimal
    # deep out of the money: the Black-Scholes put is 1e-3, so 5000 paths see
    # almost no exercise; the value is the oracle's and tiny
    bs_put = orc.bs_call(200, 100, 1, .03, .2) - 200 + 100 * np.exp(-.03)
    op, ose, oeu = orc.lsmc(P["S0"], P["K"], P["sigma"], P["r"], P["T"], P["steps"], P["paths"], seed=2)
    assert abs(p[2] - op[2]) <= 1e-12 * max(op[2], 1e-300) and p[2] >= eu[2] - 1e-12 and p[2] < 10 * bs_put
    with pytest.raises(mc.FsmcError):
        _call(engine, dict(P, steps=[0, 30, 30]), mc.mode())
    with pytest.raises(mc.FsmcError):
        engine.american_lsmc_ref(100, 10, 100.0, 110.0, .25, 2.0, mc.mode(rng=mc.RNG_BUFFER))


def _sample_stream_buffer(draw, n_paths, n_steps, loop):
    # the sample's single-thread draw: step s of path p is stream position
    # (n_paths // 1024) (s - 1) + p + loop n_paths
    nb = n_paths // 1024
    G = draw(12345, loop * n_paths, n_paths + nb * (n_steps - 1))
    return np.stack([G[nb * j:nb * j + n_paths] for j in range(n_steps)])


@pytest.mark.parametrize("loop", [0, 3])
def test_reference_algorithm_on_sample_stream(engine, loop):
    n_paths, n_steps = 8192, 64
    m = mc.mode(mc.FP64, mc.RNG_MRG32K3A, seed=12345, stream=loop)
    am, eu = engine.american_lsmc_ref(n_paths, n_steps, 100.0, 110.0, .25, 2.0, m)
    # the same normals through buffer mode: identical results
    z = _sample_stream_buffer(engine.mrg32k3a_gaussian, n_paths, n_steps, loop)
    bam, beu = engine.american_lsmc_ref(n_paths, n_steps, 100.0, 110.0, .25, 2.0, mc.mode(mc.FP64, mc.RNG_BUFFER),
                                        z)
    assert am == bam and eu == beu
    # and the oracle on its own MRG32k3a stream
    oz = _sample_stream_buffer(orc.mrg32k3a_gaussian, n_paths, n_steps, loop)
    oam, oeu = orc.lsmc_ref(n_paths, n_steps, 100.0, 110.0, .25, 2.0, oz)
    assert abs(am - oam) <= 1e-12 * abs(oam) and abs(eu - oeu) <= 1e-12 * abs(oeu)


# ---- resident sweep (one cooperative launch, state in registers) ----------
def _plan_prices(engine, P, m, sweep):
    plan = engine.lsmc_plan(P["S0"], P["K"], P["sigma"], P["r"], P["T"], P["steps"], P["paths"], m, sweep=sweep)
    try:
        got = plan.sweep
        plan.run()
        return got, plan.results()
    finally:
        plan.close()


def test_resident_fp64_matches_oracle(engine):
    # uneven sizes: parts of 4096 paths, a last part of 1 path, paths not a multiple of 4, one-step and two-step options
    P = dict(S0=[90.0, 100.0, 110.0, 95.0, 100.0, 85.0], K=[100.0] * 6, sigma=[.2, .3, .4, .15, .25, .35],
             r=[.03] * 6, T=[1.0, .5, 2.0, .25, 1.0, .75], steps=[20, 9, 13, 2, 1, 17],
             paths=[9000, 1025, 4097, 700, 3000, 8192])
    got, (p, se, eu) = _plan_prices(engine, P, mc.mode(mc.FP64, seed=21), mc.LSMC_RESIDENT)
    assert got == mc.LSMC_RESIDENT
    op, ose, oeu = orc.lsmc(P["S0"], P["K"], P["sigma"], P["r"], P["T"], P["steps"], P["paths"], seed=21)
    assert np.allclose(p, op, rtol=1e-12, atol=0)
    assert np.allclose(se, ose, rtol=1e-9, atol=0)
    assert np.allclose(eu, oeu, rtol=1e-12, atol=0)


def test_resident_long_option_reads_dates_from_memory(engine):
    # more dates than the kernel stages in LDS (1024): that option's date
    # constants are read from device memory; same results as the oracle
    P = dict(S0=[100.0, 90.0], K=[100.0] * 2, sigma=[.25, .3], r=[.03] * 2, T=[2.0, 1.0], steps=[1100, 40],
             paths=[3000, 5000])
    got, (p, se, eu) = _plan_prices(engine, P, mc.mode(mc.FP64, seed=31), mc.LSMC_RESIDENT)
    assert got == mc.LSMC_RESIDENT
    op, ose, oeu = orc.lsmc(P["S0"], P["K"], P["sigma"], P["r"], P["T"], P["steps"], P["paths"], seed=31)
    assert np.allclose(p, op, rtol=1e-12, atol=0)
    assert np.allclose(eu, oeu, rtol=1e-12, atol=0)


def test_resident_and_streamed_sweeps_agree_256_puts(engine):
    # the 256-put workload: both sweeps price the same paths; FP64 differs only
    # by the order the moment sums are combined in, FP32 by the rounding of
    # that order (a handful of exercise decisions at the regression boundary)
    P = _opts(256, 3)
    res = {}
    for prec in (mc.FP32, mc.FP64):
        for sweep in (mc.LSMC_RESIDENT, mc.LSMC_STREAMED):
            got, out = _plan_prices(engine, P, mc.mode(prec, seed=3), sweep)
            assert got == sweep
            res[prec, sweep] = out
    p_r, se_r, eu_r = res[mc.FP64, mc.LSMC_RESIDENT]
    p_s, se_s, eu_s = res[mc.FP64, mc.LSMC_STREAMED]
    assert np.allclose(p_r, p_s, rtol=1e-11, atol=0)
    assert np.allclose(eu_r, eu_s, rtol=1e-12, atol=0)
    q_r, _, qeu_r = res[mc.FP32, mc.LSMC_RESIDENT]
    q_s, _, qeu_s = res[mc.FP32, mc.LSMC_STREAMED]
    d = np.abs(q_r - q_s) / se_s
    assert np.median(d) < 0.1 and np.max(d) < 1.0, (np.median(d), np.max(d))
    assert np.allclose(qeu_r, qeu_s, rtol=1e-5, atol=0)
    # and the resident FP32 against FP64 on the same stream, as for the streamed sweep
    d64 = np.abs(q_r - p_s) / se_s
    assert np.median(d64) < 0.15 and np.max(d64) < 1.5, (np.median(d64), np.max(d64))


def test_resident_reruns_are_identical(engine):
    P = _opts(40, 6)
    m = mc.mode(mc.FP32, seed=11)
    plan = engine.lsmc_plan(P["S0"], P["K"], P["sigma"], P["r"], P["T"], P["steps"], P["paths"], m,
                            sweep=mc.LSMC_RESIDENT)
    try:
        assert plan.sweep == mc.LSMC_RESIDENT
        plan.run()
        a = plan.results()
        plan.run()
        b = plan.results()
    finally:
        plan.close()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    # the default sweep of a plan that fits is the resident one
    c = _call(engine, P, m)
    assert all(np.array_equal(x, y) for x, y in zip(a, c))


def test_resident_sweep_needs_every_option_to_fit(engine):
    # an option of more parts than the resident grid holds: AUTO streams it,
    # RESIDENT refuses before any launch
    P = dict(S0=[100.0], K=[100.0], sigma=[.2], r=[.03], T=[1.0], steps=[3], paths=[4096 * 4096])
    m = mc.mode(mc.FP32, seed=1)
    plan = engine.lsmc_plan(P["S0"], P["K"], P["sigma"], P["r"], P["T"], P["steps"], P["paths"], m)
    try:
        assert plan.sweep == mc.LSMC_STREAMED
    finally:
        plan.close()
    with pytest.raises(mc.FsmcError):
        engine.lsmc_plan(P["S0"], P["K"], P["sigma"], P["r"], P["T"], P["steps"], P["paths"], m,
                         sweep=mc.LSMC_RESIDENT)
