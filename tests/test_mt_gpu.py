"""MKL-compatible MT19937 stream on the GPU (the AsianOptionsPricing sample's
generator, mkl_rng.h:5-69, MCAsianOptions.cpp:92-93,134-135) against the CPU
oracle, whose outputs are pinned to MKL's and numpy's (test_oracle_pins.py):
raw outputs bit for bit across chained calls, Box-Muller normals to FP64
rounding, and the sample's chain (per option, each stream's next
paths x steps normals, then the buffer-mode Heston kernel) against the
oracle's Heston on the oracle's normals."""
import os
import re
import subprocess

import numpy as np
import pytest

import mcengine as mc
import oracle as orc

pytestmark = pytest.mark.gpu


def test_bits_chain_across_calls(engine):
    for seed in (42, 43, 2**32 - 1):
        g = engine.mt19937(seed)
        parts = [g.bits(n) for n in (1000, 700, 0, 5000, 624, 1)]
        g.close()
        got = np.concatenate(parts)
        assert np.array_equal(got, orc.mt19937_bits(seed, 0, got.size))
    g = engine.mt19937(42)
    assert list(g.bits(3)) == [2746317213, 478163327, 107420369]  # MKL, seed 42
    g.close()


def test_skip_ahead_calls_equal_the_sequential_stream(engine):
    # calls of at least 2 x 624 x 64 outputs are split over substreams that
    # each start by a skip-ahead (x^(kL) mod phi): the concatenated outputs
    # of a chain of small and large calls equal the oracle's sequential
    # stream bit for bit, and every call continues where the last one ended
    for seed in (42, 7):
        g = engine.mt19937(seed)
        sizes = (5, 79872, 1000, 300001, 624, 2000000, 3)
        print(f"mt19937 skip-ahead chain seed {seed}: {sizes}", flush=True)
        got = np.concatenate([g.bits(n) for n in sizes])
        g.close()
        want = orc.mt19937_bits(seed, 0, got.size)
        assert np.array_equal(got, want)
    # normals of a large call = the oracle's Box-Muller on the same stream
    g = engine.mt19937(43)
    z = g.gaussian(600000)
    g.close()
    oz = orc.mt19937_gaussian(43, 0, 600000)
    assert np.max(np.abs(z - oz) / np.maximum(np.abs(oz), 1.0)) < 1e-13


def test_gaussian_chain_and_async(engine):
    import torch
    g = engine.mt19937(42)
    a = g.gaussian(8192)
    b = g.gaussian(8192, mean=0.5, stddev=2.0)
    t = torch.empty(4096, dtype=torch.float64, device="cuda")
    g.gaussian_async(4096, t)
    torch.cuda.synchronize()
    c = t.cpu().numpy()
    g.close()
    oz = orc.mt19937_gaussian(42, 0, 8192 * 2 + 4096)
    want = np.concatenate([oz[:8192], 0.5 + 2.0 * oz[8192:16384], oz[16384:]])
    got = np.concatenate([a, b, c])
    # FP64 log / sincospi on the GPU against libm log / sin / cos
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1.0)) < 1e-13
    with pytest.raises(mc.FsmcError):
        engine.mt19937(1).gaussian(3)  # Box-Muller pairs need an even count


@pytest.mark.parametrize("prec", [mc.FP64, mc.FP32])
def test_asian_sample_chain(engine, prec):
    # the Asian command line's chain at a small size: per option, the next
    # paths x steps normals of each stream, priced in buffer mode with the
    # sample's quirks; two consecutive pricings continue the streams
    n_opt, n_paths, n_steps, r, rho = 3, 256, 16, 0.03, -0.65
    rng = np.random.default_rng(11)
    P = [rng.uniform(5, 50, n_opt), rng.uniform(10, 25, n_opt), rng.uniform(1, 5, n_opt),
         rng.uniform(.01, .015, n_opt), rng.uniform(6, 7, n_opt), rng.uniform(.02, .03, n_opt),
         rng.uniform(.45, .6, n_opt)]
    q = mc.QUIRK_ASIAN_VAR_AS_VOL | mc.QUIRK_ASIAN_RUNNING_VAL2
    m = mc.mode(prec, mc.RNG_BUFFER, quirks=q)
    g1, g2 = engine.mt19937(42), engine.mt19937(43)
    per = n_paths * n_steps
    for loop in range(2):
        for o in range(n_opt):
            z1, z2 = g1.gaussian(per), g2.gaussian(per)
            args = [[a[o]] for a in P]
            res, conf = engine.heston_asian(*args, r, rho, n_paths, n_steps, m, z1=z1, z2=z2)
            k = (loop * n_opt + o) * per
            oz1, oz2 = orc.mt19937_gaussian(42, k, per), orc.mt19937_gaussian(43, k, per)
            ores, oconf = orc.heston_buffer(*args, r, rho, n_paths, n_steps, oz1, oz2, quirks=orc.Q_VAR_AS_VOL
                                            | orc.Q_RUNNING_VAL2)
            tol = 1e-12 if prec == mc.FP64 else 2e-5
            assert res[0] == pytest.approx(ores[0], rel=tol, abs=tol)
    g1.close()
    g2.close()


@pytest.mark.parametrize("build,tol", [("dp", 1e-12), ("sp", 2e-5)])
def test_asian_command_line_prices(build, tol):
    # bin/asian_{dp,sp} at 2 options x 1024 paths x 16 steps: its parameters
    # come from the process's unseeded std::rand() (glibc, seed 1) in the
    # sample's order; the printed prices are the second pricing's, on
    # positions (n_opt + o) x paths x steps onward of both streams
    import ctypes
    n_opt, n_paths, n_steps = 2, 1024, 16
    exe = os.path.join(mc.HERE, "bin", f"asian_{build}")
    r = subprocess.run([exe, str(n_opt), str(n_paths), str(n_steps), "prices"], capture_output=True, text=True,
                       timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    assert float(re.search(r"Completed RNG\s+([0-9.]+)", r.stdout).group(1)) > 0
    prices = [float(x) for x in re.findall(r"option \d+ price ([0-9.eE+-]+)", r.stdout)]
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    draw = lambda lo, hi: (lambda t: (1.0 - t) * lo + t * hi)(libc.rand() / 2147483647.0)
    ranges = [(5.0, 50.0), (10.0, 25.0), (1.0, 5.0), (.01, .015), (6.0, 7.0), (.02, .03), (.45, .6)]
    P = np.array([[draw(lo, hi) for lo, hi in ranges] for _ in range(n_opt)])
    per = n_paths * n_steps
    for o in range(n_opt):
        k = (n_opt + o) * per
        ores, _ = orc.heston_buffer(*[[v] for v in P[o]], 0.03, -0.65, n_paths, n_steps,
                                    orc.mt19937_gaussian(42, k, per), orc.mt19937_gaussian(43, k, per),
                                    quirks=orc.Q_VAR_AS_VOL | orc.Q_RUNNING_VAL2)
        assert prices[o] == pytest.approx(ores[0], rel=tol, abs=tol)


def test_streams_against_mkl_fixture(engine):
    # MKL 2021.4's own outputs (tests/golden/mkl_streams.npz): MT19937 bits
    # bitwise and BoxMuller2 normals to MKL's arithmetic; MRG32k3a uniforms
    # bitwise and ICDF normals to MKL's inverse-CDF accuracy
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mkl_streams.npz"))
    for seed in (42, 43):
        g = engine.mt19937(seed)
        assert np.array_equal(g.bits(2048), d[f"mt{seed}_bits"])
        g.close()
        g = engine.mt19937(seed)
        assert np.max(np.abs(g.gaussian(2048) - d[f"mt{seed}_boxmuller2"])) < 3e-8
        g.close()
    assert np.array_equal(engine.mrg32k3a_uniform(12345, 0, 2048), d["mrg12345_uniform"])
    f = d["mrg12345_icdf"]
    assert np.max(np.abs(engine.mrg32k3a_gaussian(12345, 0, 2048) - f) / np.abs(f)) < 8e-9
    f = d["mrg12345_skip1000003_icdf"]
    assert np.max(np.abs(engine.mrg32k3a_gaussian(12345, 1000003, 256) - f) / np.abs(f)) < 8e-9
