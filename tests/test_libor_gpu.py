"""LIBOR swaption portfolio and adjoint deltas on the GPU against the CPU
oracle (a restatement of liborSwaptionGreeks.cpp:34-211 pinned by per-rate
finite differences).  FP64 agrees with the oracle to <= 1e-12 relative on the
same normals; FP32 errors are stated against the FP64 GPU run in units of the
standard error."""
import numpy as np
import pytest

import mcengine as mc
import oracle as orc

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


@pytest.mark.parametrize("quirks", [0, mc.QUIRK_LIBOR_INPLACE_ADJ])
def test_buffer_fp64_matches_oracle(engine, quirks):
    n = 1000
    z = np.random.default_rng(1).standard_normal((n, 40))
    v, vse, d, dse = engine.libor(n, mc.mode(mc.FP64, mc.RNG_BUFFER, quirks=quirks), L0=0.051, z=z)
    oq = orc.Q_LIBOR_INPLACE if quirks else 0
    ov, ovse, od, odse = orc.libor(n, L0=0.051, z=z, quirks=oq)
    assert abs(v - ov) <= 1e-12 * abs(ov)
    assert abs(vse - ovse) <= 1e-10 * ovse
    assert _rel(d, od) <= 1e-12
    assert _rel(dse, odse) <= 1e-9


def test_buffer_fp32_matches_oracle(engine):
    n = 2048
    z = np.random.default_rng(2).standard_normal((n, 40))
    v, vse, d, dse = engine.libor(n, mc.mode(mc.FP32, mc.RNG_BUFFER), L0=0.051, z=z)
    ov, ovse, od, odse = orc.libor(n, L0=0.051, z=z)
    assert abs(v - ov) <= 1e-5 * abs(ov)
    assert np.all(np.abs(d - od) <= 1e-5 * np.abs(od) + 1e-3 * odse)


@pytest.mark.parametrize("lam", [0.05, 0.3, 0.31, 0.45])
def test_buffer_fp32_lambda_range(engine, lam):
    # lambda <= 0.3 rebuilds the rows in registers (0.3 is the series' edge),
    # larger lambdas store the trajectory; both agree with the oracle.  (At
    # lambda = 0.6 FP32 rates overflow on some paths, and at lambda = 0 every
    # path is the same, so the mean keeps the 4e-5 FP32 rounding of the
    # swaption payoffs: neither is an FP32 workload.)
    n = 2048
    z = np.random.default_rng(3).standard_normal((n, 40))
    v, vse, d, dse = engine.libor(n, mc.mode(mc.FP32, mc.RNG_BUFFER), L0=0.051, lam=lam, z=z)
    ov, ovse, od, odse = orc.libor(n, L0=0.051, lam=lam, z=z)
    assert abs(v - ov) <= 1e-5 * abs(ov)
    assert np.all(np.abs(d - od) <= 1e-5 * np.abs(od) + 1e-3 * odse)


@pytest.mark.parametrize("n", [4096, 1001])
def test_philox_fp64_matches_oracle(engine, n):
    v, vse, d, dse = engine.libor(n, mc.mode(mc.FP64, seed=17, stream=5))
    ov, ovse, od, odse = orc.libor(n, seed=17, stream=5)
    assert abs(v - ov) <= 1e-12 * abs(ov)
    assert _rel(d, od) <= 1e-12


def test_mrg32k3a_stream_matches_oracle(engine):
    # generateRand's stream on the GPU: uniforms bit for bit, normals to the
    # last bits of the FP64 log/sqrt in the inverse CDF
    for seed, off, n in ((12345, 0, 5000), (7, 123457, 3001), (4000000000, 2**40 + 5, 2048)):
        u = engine.mrg32k3a_uniform(seed, off, n)
        assert np.array_equal(u, orc.mrg32k3a_uniform(seed, off, n))
        z = engine.mrg32k3a_gaussian(seed, off, n, 0.5, 2.0)
        oz = orc.mrg32k3a_gaussian(seed, off, n, 0.5, 2.0)
        assert np.max(np.abs(z - oz) / np.maximum(np.abs(oz), 1.0)) < 1e-14
    assert engine.mrg32k3a_gaussian(12345, 0, 0).size == 0


def test_sample_stream_mode(engine):
    # FSMC_RNG_MRG32K3A: generateRand's normals generated in place, equal to
    # buffer mode on the same stream (stream id = block of 40 n normals)
    n = 1500
    for stream in (0, 2):
        a = engine.libor(n, mc.mode(mc.FP64, mc.RNG_MRG32K3A, seed=12345, stream=stream), L0=0.051)
        z = engine.mrg32k3a_gaussian(12345, stream * 40 * n, 40 * n).reshape(n, 40)
        b = engine.libor(n, mc.mode(mc.FP64, mc.RNG_BUFFER), L0=0.051, z=z)
        assert a[0] == b[0] and np.array_equal(a[2], b[2])
    oz = orc.mrg32k3a_gaussian(12345, 0, 40 * n).reshape(n, 40)
    ov, _, od, _ = orc.libor(n, L0=0.051, z=oz)
    a = engine.libor(n, mc.mode(mc.FP64, mc.RNG_MRG32K3A, seed=12345), L0=0.051)
    assert abs(a[0] - ov) <= 1e-12 * abs(ov) and _rel(a[2], od) <= 1e-12
    with pytest.raises(mc.FsmcError):  # not for the Philox-only workloads
        engine.gbm_european([10.0], [10.0], [1.0], .02, .3, 100, mc.mode(rng=mc.RNG_MRG32K3A))


def test_normals_seam_equals_philox(engine):
    # generateRand replacement: buffer mode on the filled z == Philox mode
    n = 512
    m = mc.mode(mc.FP64, seed=23)
    z = engine.libor_normals(n, m)
    a = engine.libor(n, m)
    b = engine.libor(n, mc.mode(mc.FP64, mc.RNG_BUFFER), z=z)
    assert abs(a[0] - b[0]) <= 1e-14 * abs(a[0])
    assert np.allclose(a[2], b[2], rtol=1e-14, atol=0)
    zs = engine.libor_normals(n, m, mean=1.0, stddev=2.0)
    assert np.allclose(zs, 1.0 + 2.0 * z, rtol=1e-15, atol=1e-15)


def test_portfolio_262144_paths_fp32_vs_fp64(engine):
    # 262144 paths, L0 = 0.05, lambda = 0.2, FP32 vs FP64 on the
    # same Philox stream; sum of deltas vs a flat-curve central difference.
    n = 262144
    v32, s32, d32, ds32 = engine.libor(n, mc.mode(mc.FP32, seed=1), L0=0.05)
    v64, s64, d64, ds64 = engine.libor(n, mc.mode(mc.FP64, seed=1), L0=0.05)
    assert abs(v32 - v64) < 0.01 * s64
    assert np.all(np.abs(d32 - d64) < 0.01 * ds64)
    h = 1e-6
    up = engine.libor(n, mc.mode(mc.FP64, seed=1), L0=0.05 + h)[0]
    dn = engine.libor(n, mc.mode(mc.FP64, seed=1), L0=0.05 - h)[0]
    assert np.sum(d64) == pytest.approx((up - dn) / (2 * h), rel=1e-5)


@pytest.mark.parametrize("lam", [0.2, 0.3])
def test_fp32_checkpointed_every_8_steps(engine, lam):
    # checkpoint_every = 8: rows 8, 16, 24, 32 stored by the forward sweep and
    # each segment rebuilt from the exact row at its top.  Against the
    # oracle on the same normals (FP64 reference arithmetic) and against the
    # kernel without checkpoints, which differs only by the inversion's
    # truncation carried across segments
    n = 2048
    z = np.random.default_rng(2).standard_normal((n, 40))
    m = mc.mode(mc.FP32, mc.RNG_BUFFER)
    v, vse, d, dse = engine.libor(n, m, L0=0.051, lam=lam, z=z, checkpoint_every=8)
    ov, ovse, od, odse = orc.libor(n, L0=0.051, lam=lam, z=z)
    assert abs(v - ov) <= 1e-5 * abs(ov)
    assert np.all(np.abs(d - od) <= 1e-5 * np.abs(od) + 1e-3 * odse)
    v0, _, d0, _ = engine.libor(n, m, L0=0.051, lam=lam, z=z)
    assert v == v0  # the value is the forward's, untouched by the reverse sweep
    assert np.all(np.abs(d - d0) <= 2e-6 * np.abs(od) + 1e-4 * odse)
    # and at full size on the Philox stream against FP64
    n = 262144
    v32, _, d32, _ = engine.libor(n, mc.mode(mc.FP32, seed=1), L0=0.05, lam=lam, checkpoint_every=8)
    v64, s64, d64, ds64 = engine.libor(n, mc.mode(mc.FP64, seed=1), L0=0.05, lam=lam)
    assert abs(v32 - v64) < 0.01 * s64
    assert np.all(np.abs(d32 - d64) < 0.01 * ds64)
    with pytest.raises(mc.FsmcError):
        engine.libor(64, m, z=z[:64], checkpoint_every=4)


def test_deterministic_and_async(engine):
    import torch
    n = 3000
    m = mc.mode(mc.FP32, seed=3)
    a = engine.libor(n, m)
    b = engine.libor(n, m)
    assert a[0] == b[0] and np.array_equal(a[2], b[2])
    dev = torch.device("cuda:0")
    ov = torch.zeros(2, dtype=torch.float64, device=dev)
    od = torch.zeros(160, dtype=torch.float64, device=dev)
    ws = torch.empty(mc.libor_workspace(n, m), dtype=torch.uint8, device=dev)
    mc.libor_async(n, m, ov, od, ws)
    torch.cuda.synchronize()
    assert ov[0].item() == a[0] and np.array_equal(od[:80].cpu().numpy(), a[2])


def test_errors(engine):
    with pytest.raises(mc.FsmcError):
        engine.libor(1, mc.mode())
    with pytest.raises(mc.FsmcError):
        engine.libor(100, mc.mode(rng=mc.RNG_BUFFER))


@pytest.mark.parametrize("build,rel", [("dp", 1e-5), ("sp", 1e-4)])
def test_sample_command_line(build, rel):
    # bin/libor_{dp,sp}: the sample's main over the drop-in liborSwaptionGreeks;
    # its printed value/lambda are 10x the per-path means (NLOOPS) on the
    # generateRand normals (MRG32k3a seed 12345, inverse CDF, stream order),
    # L0 = 0.051, lambda = 0.2.
    import os
    import re
    import subprocess
    exe = os.path.join(mc.HERE, "bin", f"libor_{build}")
    r = subprocess.run([exe, "2048"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    m = re.search(r"Result: value=([-0-9.eE+]+),\s*lambda=([-0-9.eE+]+)", r.stdout)
    assert m and re.search(r"Total time: [0-9.eE+]+ ms", r.stdout), r.stdout
    z = orc.mrg32k3a_gaussian(12345, 0, 2048 * 40).reshape(2048, 40)
    ov, _, od, _ = orc.libor(2048, L0=0.051, z=z)
    assert float(m.group(1)) == pytest.approx(10 * ov, rel=rel)
    assert float(m.group(2)) == pytest.approx(10 * od[79], rel=rel)
