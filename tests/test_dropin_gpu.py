"""The C++ drop-ins of the European and American samples on the GPU, called
through their mangled C++ symbols as a sample's object files call them.

perform_sims (utilities.cpp:42-88) fills the caller's rows with the path
matrix on the sample's own MRG32k3a stream layout; it is checked against the
oracle's MRG32k3a + inverse-CDF normals pushed through the sample's
recursion in numpy (FP64 to 1e-11 relative: the normals agree to ~1e-15 and
exp2 to an ulp per step), FP32 against FP64 to 2e-5.  price_american_lsmc and
price_european must equal the C-ABI calls they wrap bit for bit."""
import ctypes
import os

import numpy as np
import pytest

import mcengine as mc
import oracle as orc

pytestmark = pytest.mark.gpu

LOG2E = 1.4426950408889634


class AmericanConfig(ctypes.Structure):
    _fields_ = [("num_paths", ctypes.c_int), ("num_steps", ctypes.c_int), ("seed", ctypes.c_uint),
                ("pricing", ctypes.c_uint)]


def _cfg():
    fn = mc.lib()._ZN11fsmc_dropin15american_configEv
    fn.restype = ctypes.POINTER(AmericanConfig)
    return fn().contents


def _sample_paths(n_paths, n_steps, spot, fwd, vs, loop):
    # utilities.cpp:52-84 single-threaded: step s of path p takes stream
    # position (n_paths // 1024) (s - 1) + p + loop n_paths
    nb = n_paths // 1024
    g = orc.mrg32k3a_gaussian(12345, loop * n_paths, n_paths + nb * (n_steps - 1))
    a = np.empty((n_steps + 1, n_paths))
    a[0] = spot
    for s in range(1, n_steps + 1):
        a[s] = a[s - 1] * np.exp2(fwd + vs * g[nb * (s - 1):nb * (s - 1) + n_paths])
    return a


def _fwd_vs(vol=0.25, maturity=2.0, steps=256):
    # main.cpp:55-59
    dt = maturity / steps
    return -0.5 * vol * vol * dt * LOG2E, vol * np.sqrt(dt) * LOG2E


@pytest.mark.parametrize("n_paths,n_steps,loop", [(4096, 32, 0), (5120, 16, 3), (1300, 8, 1)])
def test_paths_ref_on_sample_stream(engine, n_paths, n_steps, loop):
    fwd, vs = _fwd_vs(steps=n_steps)
    want = _sample_paths(n_paths, n_steps, 100.0, fwd, vs, loop)
    m = mc.mode(mc.FP64, mc.RNG_MRG32K3A, seed=12345, stream=loop)
    got = engine.american_paths_ref(n_paths, n_steps, 100.0, fwd, vs, m)
    assert np.max(np.abs(got / want - 1)) < 1e-11
    # the same normals through buffer mode give the same matrix bit for bit
    nb = n_paths // 1024
    z = engine.mrg32k3a_gaussian(12345, loop * n_paths, n_paths + nb * (n_steps - 1))
    zb = np.stack([z[nb * j:nb * j + n_paths] for j in range(n_steps)])
    assert np.array_equal(engine.american_paths_ref(n_paths, n_steps, 100.0, fwd, vs, mc.mode(mc.FP64, mc.RNG_BUFFER),
                                                    zb), got)
    g32 = engine.american_paths_ref(n_paths, n_steps, 100.0, fwd, vs, mc.mode(mc.FP32, mc.RNG_MRG32K3A, seed=12345,
                                                                                stream=loop))
    assert g32.dtype == np.float32 and np.max(np.abs(g32 / got - 1)) < 2e-5


def test_paths_ref_matches_lsmc_ref_sweep(engine):
    # the matrix perform_sims returns drives the sample's sweep: the oracle's
    # sweep on the oracle's normals equals the GPU's full LSMC on the same stream
    n_paths, n_steps = 4096, 24
    m = mc.mode(mc.FP64, mc.RNG_MRG32K3A, seed=12345, stream=2)
    am, eu = engine.american_lsmc_ref(n_paths, n_steps, 100.0, 110.0, .25, 2.0, m)
    nb = n_paths // 1024
    g = orc.mrg32k3a_gaussian(12345, 2 * n_paths, n_paths + nb * (n_steps - 1))
    oam, oeu = orc.lsmc_ref(n_paths, n_steps, 100.0, 110.0, .25, 2.0,
                            np.stack([g[nb * j:nb * j + n_paths] for j in range(n_steps)]))
    assert abs(am - oam) <= 1e-12 * abs(oam) and abs(eu - oeu) <= 1e-12 * abs(oeu)


@pytest.mark.parametrize("real", [np.float64, np.float32])
def test_perform_sims_dropin(engine, real):
    cfg = _cfg()
    saved = (cfg.num_paths, cfg.num_steps, cfg.pricing)
    n_paths, n_steps = 2048, 12
    cfg.num_paths, cfg.num_steps, cfg.pricing = n_paths, n_steps, 0
    try:
        ct = ctypes.c_double if real == np.float64 else ctypes.c_float
        sym = "_Z12perform_simsPPdPPvidd" if real == np.float64 else "_Z12perform_simsPPfPPviff"
        fn = getattr(mc.lib(), sym)
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ct, ct]
        fn.restype = None
        fwd, vs = _fwd_vs(steps=n_steps)
        prec = mc.FP64 if real == np.float64 else mc.FP32
        # the sample's main calls it once per loop on one stream table
        # (main.cpp:117-121): loop l reads run l of every stream
        for loop in range(2):
            a = np.full((n_steps + 1, n_paths), 100.0, dtype=real)
            rows = (ctypes.c_void_p * (n_steps + 1))(*[a[i].ctypes.data for i in range(n_steps + 1)])
            fn(rows, None, 8, fwd, vs)
            assert cfg.pricing == loop + 1
            want = engine.american_paths_ref(n_paths, n_steps, 100.0, float(real(fwd)), float(real(vs)),
                                             mc.mode(prec, mc.RNG_MRG32K3A, seed=12345, stream=loop))
            assert np.array_equal(a, want)
            assert np.all(a[0] == 100.0)
    finally:
        cfg.num_paths, cfg.num_steps, cfg.pricing = saved


@pytest.mark.parametrize("real", [np.float64, np.float32])
def test_price_american_lsmc_dropin(engine, real):
    ct = ctypes.c_double if real == np.float64 else ctypes.c_float
    sym = "_Z19price_american_lsmciiddddPdS_j" if real == np.float64 else "_Z19price_american_lsmciiffffPfS_j"
    fn = getattr(mc.lib(), sym)
    fn.argtypes = [ctypes.c_int, ctypes.c_int] + [ct] * 4 + [ctypes.POINTER(ct)] * 2 + [ctypes.c_uint]
    fn.restype = None
    am, eu = ct(), ct()
    fn(8192, 32, 100.0, 110.0, 0.25, 2.0, ctypes.byref(am), ctypes.byref(eu), 1)
    prec = mc.FP64 if real == np.float64 else mc.FP32
    wam, weu = engine.american_lsmc_ref(8192, 32, 100.0, 110.0, .25, 2.0,
                                        mc.mode(prec, mc.RNG_MRG32K3A, seed=12345, stream=1))
    assert am.value == real(wam) and eu.value == real(weu)
    # no rate in the sample's dynamics (main.cpp:55-59): early exercise is
    # worth nothing, so the American price is the European one less the
    # regression's low bias
    assert 0.95 * eu.value < am.value < 1.01 * eu.value and 15.0 < eu.value < 25.0


class EuropeanConfig(ctypes.Structure):
    _fields_ = [("sample_stream", ctypes.c_bool), ("threads", ctypes.c_int)]


def _european_cfg():
    fn = mc.lib()._ZN11fsmc_dropin15european_configEv
    fn.restype = ctypes.POINTER(EuropeanConfig)
    return fn().contents


def _price_european(real, S, X, T, r, sig, paths, block, seed):
    ct = ctypes.c_double if real == np.float64 else ctypes.c_float
    sym = "_Z14price_europeanPKdS0_S0_ddiiiPdS1_y" if real == np.float64 else "_Z14price_europeanPKfS0_S0_ffiiiPfS1_y"
    fn = getattr(mc.lib(), sym)
    P = ctypes.POINTER(ct)
    fn.argtypes = [P, P, P, ct, ct, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_ulonglong]
    fn.restype = None
    S, X, T = (np.ascontiguousarray(v, dtype=real) for v in (S, X, T))
    res, conf = np.zeros(len(S), dtype=real), np.zeros(len(S), dtype=real)
    fn(*[v.ctypes.data_as(P) for v in (S, X, T)], r, sig, len(S), paths, block, res.ctypes.data_as(P),
       conf.ctypes.data_as(P), seed)
    return res, conf


def test_price_european_sample_stream(engine):
    # the sample's own normals: thread t of 2 prices options [2t, 2t + 2) from
    # SFMT19937 seed 777 + t, paths in order, s = MuByT + VBySqrtT z (float);
    # the oracle's SFMT19937 (pinned to MKL's outputs) through the oracle's
    # buffer-mode estimator
    cfg = _european_cfg()
    saved = (cfg.sample_stream, cfg.threads)
    cfg.sample_stream, cfg.threads = True, 2
    try:
        from scipy.special import ndtri
        rng = np.random.default_rng(3)
        S, X, T = rng.uniform(5, 50, 4), rng.uniform(10, 25, 4), rng.uniform(1, 5, 4)
        r, sig, paths = 0.06, 0.10, 1024
        res, conf = _price_european(np.float64, S, X, T, r, sig, paths, 256, 777)
        z = np.concatenate([ndtri(orc.sfmt19937_icdf_uniforms(777 + t, 2 * paths)) for t in range(2)])
        z = z.reshape(4, paths)
        mu = (LOG2E * (r - .5 * sig * sig) * T)[:, None]
        vol = (LOG2E * sig * np.sqrt(T))[:, None]
        samples = (mu + vol * z).astype(np.float32)
        ores, oconf = orc.gbm_buffer(S, X, T, r, sig, samples)
        assert np.max(np.abs(res - ores) / np.maximum(ores, 1e-12)) < 1e-6
        assert np.max(np.abs(conf - oconf) / np.maximum(oconf, 1e-12)) < 1e-5
        cfg.threads = 3  # 4 options cannot be split over 3 sample threads
        bad, _ = _price_european(np.float64, S, X, T, r, sig, paths, 256, 777)
        assert np.all(np.isnan(bad))
    finally:
        cfg.sample_stream, cfg.threads = saved


def test_sfmt19937_generator(engine):
    # one stream against MKL's own float normals and the oracle's exact
    # inverse CDF; several streams with per-segment mean and scale
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mkl_streams.npz"))
    from scipy.special import ndtri
    z = engine.sfmt19937_gaussian(777, 2048)
    assert np.max(np.abs(z - d["sfmt777_icdf_float"])) < 2e-5
    exact = ndtri(orc.sfmt19937_icdf_uniforms(777, 2048))
    assert np.max(np.abs(z - exact) / np.maximum(np.abs(exact), 1e-3)) < 2e-7
    import torch
    dev = torch.device("cuda:0")
    n_streams, per, seg = 3, 1500, 500  # 1500 = 2 x 624 + 252: a partial block
    mean = torch.tensor([0.5 * k for k in range(9)], dtype=torch.float64, device=dev)
    sd = torch.tensor([1.0 + k for k in range(9)], dtype=torch.float64, device=dev)
    out = torch.zeros(n_streams * per, dtype=torch.float32, device=dev)
    mc.sfmt19937_gaussian_async(5, n_streams, per, out, seg, mean, sd)
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(n_streams, per)
    for t in range(n_streams):
        e = ndtri(orc.sfmt19937_icdf_uniforms(5 + t, per))
        k = (t * per + np.arange(per)) // seg
        want = mean.cpu().numpy()[k] + sd.cpu().numpy()[k] * e
        assert np.max(np.abs(got[t] - want) / np.maximum(np.abs(want), 1e-3)) < 2e-7


def test_sfmt19937_skip_ahead_equals_sequential(engine):
    # the skip-ahead path splits each stream into K = B0 B1 substreams whose
    # start states come from jumps (round 0: x^(m B1 L / 4) mod phi from the
    # seeded state; round 1: x^(j L / 4) from those): its float normals are
    # bit for bit those of the sequential kernel, and the first ones those
    # of the oracle's SFMT19937 stream.  Cases: K = 32 (one round), 16 on
    # three streams, 128 = 16 x 8 with segments shorter than a block,
    # 512 = 32 x 16 on two streams, 1024 = 32 x 32
    import torch
    from scipy.special import ndtri
    dev = torch.device("cuda:0")
    for n_streams, per, seg in ((1, 2_000_003, 2_000_003), (3, 700_000, 350_000), (1, 4_999_800, 300),
                                (2, 21_000_000, 1_000_000), (1, 41_000_000, 41_000_000)):
        print(f"sfmt19937 skip-ahead: {n_streams} streams x {per}", flush=True)
        nseg = n_streams * per // seg
        mean = torch.linspace(-0.5, 0.5, nseg, dtype=torch.float64, device=dev)
        sd = torch.linspace(1.0, 2.0, nseg, dtype=torch.float64, device=dev)
        seq = torch.zeros(n_streams * per, dtype=torch.float32, device=dev)
        mc.sfmt19937_gaussian_async(777, n_streams, per, seq, seg, mean, sd)
        ws_bytes = mc.sfmt19937_workspace(n_streams, per)
        assert ws_bytes > 0
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        skip = torch.full((n_streams * per,), float("nan"), dtype=torch.float32, device=dev)
        mc.sfmt19937_gaussian_skip_async(777, n_streams, per, skip, ws, seg, mean, sd)
        torch.cuda.synchronize()
        a, b = seq.cpu().numpy(), skip.cpu().numpy()
        assert np.array_equal(a, b)
        e = ndtri(orc.sfmt19937_icdf_uniforms(777, 20000))
        k = np.arange(20000) // seg
        want = mean.cpu().numpy()[k] + sd.cpu().numpy()[k] * e
        assert np.max(np.abs(b[:20000] - want) / np.maximum(np.abs(want), 1e-3)) < 2e-7


def test_sfmt19937_sample_scaled_draws_against_mkl(engine):
    # the European sample's draws, vsRngGaussian(ICDF, ..., MuByT, VBySqrtT)
    # on stream 777 in three consecutive blocks, against MKL's own output
    # (tests/golden/mkl_streams.npz): MKL applies mean and deviation in float
    # (pinned on the CPU), so the gap is MKL's float inverse CDF, up to
    # 6e-5 |z| of the deviation in the tails
    import torch
    from scipy.special import ndtri
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mkl_streams.npz"))
    blocks, mean, sd = d["sfmt777_sample_blocks"], d["sfmt777_sample_mean"], d["sfmt777_sample_sd"]
    dev = torch.device("cuda:0")
    out = torch.zeros(blocks.size, dtype=torch.float32, device=dev)
    mc.sfmt19937_gaussian_async(777, 1, blocks.size, out, blocks.shape[1], torch.tensor(mean, device=dev),
                                torch.tensor(sd, device=dev))
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(blocks.shape)
    z = ndtri(orc.sfmt19937_icdf_uniforms(777, blocks.size)).reshape(blocks.shape)
    assert np.all(np.abs(got - blocks) <= sd[:, None] * 6e-5 * np.maximum(1.0, np.abs(z)) + 2e-7)
    # and the engine's value is the exact mean + sd z rounded once to float
    exact = mean[:, None] + sd[:, None] * z
    assert np.all(np.abs(got - exact) <= 1.5 * np.spacing(np.abs(exact).astype(np.float32)))


@pytest.mark.parametrize("real", [np.float64, np.float32])
def test_price_european_dropin(engine, real):
    cfg = _european_cfg()
    saved = (cfg.sample_stream, cfg.threads)
    cfg.sample_stream = False  # the engine's Philox normals
    try:
        _price_european_philox(engine, real)
    finally:
        cfg.sample_stream, cfg.threads = saved


def _price_european_philox(engine, real):
    ct = ctypes.c_double if real == np.float64 else ctypes.c_float
    sym = "_Z14price_europeanPKdS0_S0_ddiiiPdS1_y" if real == np.float64 else "_Z14price_europeanPKfS0_S0_ffiiiPfS1_y"
    fn = getattr(mc.lib(), sym)
    P = ctypes.POINTER(ct)
    fn.argtypes = [P, P, P, ct, ct, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_ulonglong]
    fn.restype = None
    rng = np.random.default_rng(7)
    n, paths = 64, 65536
    S, X, T = rng.uniform(5, 50, n), rng.uniform(10, 25, n), rng.uniform(1, 5, n)
    S, X, T = (v.astype(real) for v in (S, X, T))
    res, conf = np.zeros(n, dtype=real), np.zeros(n, dtype=real)
    fn(*[v.ctypes.data_as(P) for v in (S, X, T)], 0.06, 0.10, n, paths, 1024, res.ctypes.data_as(P),
       conf.ctypes.data_as(P), 777)
    prec = mc.FP64 if real == np.float64 else mc.FP32
    wr, wc = engine.gbm_european(S.astype(np.float64), X.astype(np.float64), T.astype(np.float64), float(real(0.06)),
                                 float(real(0.10)), paths, mc.mode(prec, mc.RNG_PHILOX, seed=777))
    assert np.array_equal(res, wr.astype(real)) and np.array_equal(conf, wc.astype(real))
    bs = np.array([orc.bs_call(float(a), float(b), float(c), 0.06, 0.10) for a, b, c in zip(S, X, T)])
    assert np.all(np.abs(res - bs) <= 5 * np.maximum(conf, 1e-6) / 1.96 + 1e-4)
