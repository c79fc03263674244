"""Closed-form Black-Scholes and the CRR lattice on the GPU against the CPU
oracle (Black-Scholes pinned by textbook values; the lattice restatement of
binomial_cpu.cpp:62-101 pinned by convergence to Black-Scholes and by the
Longstaff-Schwartz table of American puts)."""
import subprocess
import os
import re

import numpy as np
import pytest

import mcengine as mc
import oracle as orc

pytestmark = pytest.mark.gpu


def _options(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(5, 30, n), rng.uniform(1, 100, n), rng.uniform(.25, 10, n), rng.uniform(.1, .5, n)


def test_black_scholes_fp64_matches_oracle(engine):
    S, X, T, V = _options(4000, 1)
    call, put, g = engine.black_scholes(S, X, T, V, .02, mc.mode(mc.FP64), greeks=True)
    ref = np.array([orc.bs_call(*a, .02, v) for *a, v in zip(S, X, T, V)])
    # small prices are differences of O(S) terms: relative error there is
    # absolute rounding (erfc on the GPU and in libm differ in the last bit)
    ok = ref > 1e-3
    assert np.max(np.abs(call - ref)[ok] / ref[ok]) < 1e-11
    assert np.max(np.abs(call - ref)) < 1e-12 * np.max(S)
    # put-call parity
    assert np.allclose(call - put, S - X * np.exp(-.02 * T), rtol=0, atol=1e-11)
    og = np.array([orc.bs_call_greeks(*a, .02, v) for *a, v in zip(S, X, T, V)])
    assert np.allclose(g, og, rtol=1e-11, atol=1e-12)


def _cnd_poly(x):
    # Abramowitz & Stegun 26.2.17 as the sample evaluates it
    ax = np.abs(x)
    pdf = 0.39894228040143270286 * np.exp(-0.5 * ax * ax)
    k = 1.0 / (1.0 + 0.2316419 * ax)
    poly = k * (0.319381530 + k * (-0.356563782 + k * (1.781477937 + k * (-1.821255978 + k * 1.330274429))))
    c = 1.0 - pdf * poly
    return np.where(x < 0, 1.0 - c, c)


def test_black_scholes_sample_cnd(engine):
    S, X, T, V = _options(2000, 2)
    call, _ = engine.black_scholes(S, X, T, V, .02, mc.mode(mc.FP64, quirks=mc.QUIRK_BS_POLY_CND))
    sq = V * np.sqrt(T)
    d1 = (np.log(S / X) + (.02 + .5 * V * V) * T) / sq
    ref = S * _cnd_poly(d1) - X * np.exp(-.02 * T) * _cnd_poly(d1 - sq)
    assert np.max(np.abs(call - ref)) < 1e-12 * np.max(S)
    exact, _ = engine.black_scholes(S, X, T, V, .02, mc.mode(mc.FP64))
    # the polynomial's own error, below 7.5e-8 on each of the two CDF values
    assert np.all(np.abs(call - exact) < 7.5e-8 * (S + X))


def test_black_scholes_fp32(engine):
    S, X, T, V = _options(3000, 3)
    c32, _ = engine.black_scholes(S, X, T, V, .02, mc.mode(mc.FP32))
    c64, _ = engine.black_scholes(S, X, T, V, .02, mc.mode(mc.FP64))
    assert np.max(np.abs(c32 - c64)) < 2e-5 * np.max(S)


@pytest.mark.parametrize("steps", [1, 2, 17, 2048])
def test_lattice_matches_oracle(engine, steps):
    S, X, T, V = _options(64, 4 + steps)
    m = mc.mode(mc.FP64)
    eu = engine.lattice(S, X, T, V, .03, steps, m)
    oeu = np.array([orc.crr_european_call(*a, .03, v, steps) for *a, v in zip(S, X, T, V)])
    assert np.max(np.abs(eu - oeu) / np.maximum(oeu, 1e-3)) < 1e-11
    am = engine.lattice(S, X, T, V, .03, steps, m, american=True, put=True)
    oam = np.array([orc.crr_american_put(*a, .03, v, steps) for *a, v in zip(S, X, T, V)])
    assert np.max(np.abs(am - oam) / np.maximum(oam, 1e-3)) < 1e-11
    # discrete put-call parity on the European lattice; American call = European call
    ep = engine.lattice(S, X, T, V, .03, steps, m, put=True)
    assert np.allclose(eu - ep, S - X * np.exp(-.03 * T), rtol=0, atol=1e-9)
    ac = engine.lattice(S, X, T, V, .03, steps, m, american=True)
    assert np.allclose(ac, eu, rtol=1e-12, atol=1e-12)
    assert np.all(am >= ep - 1e-12)


def test_lattice_longstaff_schwartz_table(engine):
    # American puts of Longstaff & Schwartz (2001), Table 1 (finite-difference
    # values, themselves within ~1e-2 of converged lattices): S0 = 36, 40, 44;
    # K = 40, r = .06, sigma = .2, T = 1
    m = mc.mode(mc.FP64)
    am = engine.lattice([36.0, 40.0, 44.0], 40.0, 1.0, .2, .06, 8000, m, american=True, put=True)
    assert np.allclose(am, [4.478, 2.314, 1.110], atol=1e-2)
    oam = [orc.crr_american_put(s, 40.0, 1.0, .06, .2, 8000) for s in (36.0, 40.0, 44.0)]
    assert np.allclose(am, oam, rtol=1e-10, atol=0)
    a32 = engine.lattice([36.0, 40.0, 44.0], 40.0, 1.0, .2, .06, 4000, mc.mode(mc.FP32), american=True, put=True)
    assert np.allclose(a32, am, atol=5e-3)


def test_analytic_errors(engine):
    with pytest.raises(mc.FsmcError):
        engine.lattice([10.0], 10.0, 1.0, .2, .03, 0, mc.mode(mc.FP64))
    with pytest.raises(mc.FsmcError):
        engine.lattice([10.0], 10.0, 1.0, .2, .03, 8001, mc.mode(mc.FP64))
    with pytest.raises(mc.FsmcError):
        engine.black_scholes([10.0], 10.0, 1.0, .2, .03, mc.mode(quirks=mc.QUIRK_ASIAN_VAR_AS_VOL))


def _run(name, *args):
    exe = os.path.join(mc.HERE, "bin", name)
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("build", ["dp", "sp"])
def test_blackscholes_command_line(build):
    out = _run(f"blackscholes_{build}", "1")
    line = [l for l in out.splitlines() if "Elapsed" in l][0]  # MPTest: grep Elapsed, $4
    assert float(line.split()[3]) > 0
    assert re.search(r"Throughput\s+= [0-9.]+ GOptions/sec", out)
    assert "TEST PASSED" in out


@pytest.mark.parametrize("build", ["dp", "sp"])
def test_binomial_command_line(build):
    out = _run(f"binomial_{build}", "1")
    line = [l for l in out.splitlines() if "Completed" in l][0]  # MPTest: grep Completed, $3
    assert float(line.split()[2]) > 0
    assert "TEST PASSED" in out
