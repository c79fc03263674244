"""Mixed request batches (micro-batches of grouped GBM / Heston grids).

A request priced inside a batch uses the normals and per-path arithmetic of
the standalone entry point with the request index as option index, so FP64
batch results equal the standalone GPU results (and therefore the oracle) to
summation order; reruns, with eager work in between, reproduce the first run
bitwise."""
import ctypes

import numpy as np
import pytest

import mcengine as mc
import oracle as orc

pytestmark = pytest.mark.gpu


def _fields(r, k):
    return np.array([getattr(x, k) for x in r])


def test_batch_matches_standalone_heston(engine):
    n, paths, steps = 6, 3000, 37
    rng = np.random.default_rng(1)
    S = rng.uniform(80, 120, n)
    reqs = [mc.heston_request(paths, steps, S[i], 100.0, 1.0, .03, .04, 2.0, .04, .3, -.7) for i in range(n)]
    m = mc.mode(mc.FP64, seed=11)
    b = mc.Batch(engine, reqs, m, tuning=(1024, 0))
    b.run()
    out = b.results()
    b.close()
    one = np.ones(n)
    # the batch prices Heston requests with the tangent path: equal to the
    # standalone tangent entry point to summation order, to the adjoint's
    # Greeks to rounding
    import torch
    dev = torch.device("cuda:0")
    prm = [torch.tensor(v, dtype=torch.float64, device=dev) for v in (S, 100 * one, one, .04 * one, 2 * one,
                                                                      .04 * one, .3 * one)]
    tr, tc, tg = (torch.zeros(n, dtype=torch.float64, device=dev), torch.zeros(n, dtype=torch.float64, device=dev),
                  torch.zeros(n, 6, dtype=torch.float64, device=dev))
    ws = torch.empty(mc.heston_asian_tangent_workspace(n, paths, steps, .03, -.7, m), dtype=torch.uint8, device=dev)
    mc.heston_asian_tangent_async(prm, paths, steps, .03, -.7, m, tr, tc, tg, ws)
    torch.cuda.synchronize()
    tr, tg = tr.cpu().numpy(), tg.cpu().numpy()
    assert np.allclose(out["price"], tr, rtol=1e-12, atol=0)
    for k, name in enumerate(("delta", "vega", "rho")):
        assert np.allclose(out[name], tg[:, k], rtol=1e-12, atol=1e-14)
    res, conf, g = engine.heston_asian(S, 100 * one, one, .04 * one, 2 * one, .04 * one, .3 * one, .03, -.7, paths,
                                       steps, m, greeks=True)
    assert np.allclose(out["price"], res, rtol=1e-12, atol=0)
    for k, name in enumerate(("delta", "vega", "rho")):
        assert np.allclose(out[name], g[:, k], rtol=1e-9, atol=1e-14)
    # and the oracle's tangent-mode Greeks
    ores, oconf, og = orc.heston_philox(S, 100 * one, one, .04 * one, 2 * one, .04 * one, .3 * one, .03, -.7, paths,
                                        steps, seed=11, greeks=True)
    assert np.allclose(out["price"], ores, rtol=1e-12, atol=0)
    for k, name in enumerate(("delta", "vega", "rho")):
        assert np.allclose(out[name], og[:, k], rtol=1e-11, atol=1e-14)


def test_batch_matches_standalone_gbm(engine):
    n, paths = 5, 5000
    rng = np.random.default_rng(2)
    S, X, T = rng.uniform(20, 30, n), rng.uniform(5, 20, n), rng.uniform(.5, 3, n)
    reqs = [mc.gbm_request(paths, S[i], X[i], T[i], .02, .3) for i in range(n)]
    m = mc.mode(mc.FP64, seed=12)
    b = mc.Batch(engine, reqs, m, tuning=(1024, 2))
    b.run()
    out = b.results()
    b.close()
    res, conf, g = engine.gbm_european(S, X, T, .02, .3, paths, m, greeks=True)
    assert np.allclose(out["price"], res, rtol=1e-12, atol=0)
    assert np.allclose(out["price_se"] * 1.96, conf, rtol=1e-9, atol=0)
    for k, name in enumerate(("delta", "vega", "rho")):
        assert np.allclose(out[name], g[:, k], rtol=1e-12, atol=1e-14)


def test_mixed_batch_reruns_and_tuning(engine):
    import torch
    reqs = mc.mixed_requests(96, seed=5)
    m = mc.mode(mc.FP32, seed=3)
    b = mc.Batch(engine, reqs, m, tuning=(4096, 32))
    b.run()
    first = b.results()
    for _ in range(3):
        x = torch.randn(1 << 20, device="cuda")  # eager work between runs
        torch.cuda.synchronize()
        del x
        b.run()
        again = b.results()
        for k in ("price", "delta", "vega", "rho"):
            assert np.array_equal(first[k], again[k])
    assert np.all(first["latency_us"] > 0)
    # micro-batches complete in order
    assert np.all(np.diff(first["latency_us"][::32]) >= 0)
    # a different tuning changes only the summation order
    b2 = mc.Batch(engine, reqs, m, tuning=(2048, 0))
    b2.run()
    other = b2.results()
    b2.close()
    assert np.allclose(other["price"], first["price"], rtol=1e-9, atol=1e-12)
    chosen = b.autotune()
    assert chosen[0] % 1024 == 0 and chosen[1] >= 0 and chosen[2] in (0, 1) and chosen[3] in (0, 1, 2, 3, 4, 6, 8)
    assert chosen[4] in (0, 1) and (chosen[4] == 0 or chosen[3] >= 1)
    assert b.tuning() == chosen
    b.run()
    tuned = b.results()
    assert np.allclose(tuned["price"], first["price"], rtol=1e-9, atol=1e-12)
    # tuned at a fixed micro-batch size: the size is kept, requests still complete in order
    micro = b.autotune(requests_per_launch=32)
    assert micro[1] == 32 and b.tuning() == micro
    # the shape's cached tuning stays the throughput choice
    b3 = mc.Batch(engine, reqs, m)
    assert b3.tuning() == chosen
    b3.close()
    b.run()
    mt = b.results()
    b.close()
    assert np.allclose(mt["price"], first["price"], rtol=1e-9, atol=1e-12)
    assert np.all(np.diff(mt["latency_us"][::32]) >= 0)


def test_results_wait_for_a_run_on_another_stream(engine):
    # run on a torch side stream, read results on the context's own stream:
    # fsmc_batch_results waits for the run's completion event first
    import torch
    reqs = mc.mixed_requests(64, seed=8)
    b = mc.Batch(engine, reqs, mc.mode(mc.FP32, seed=4), tuning=(4096, 0))
    b.run()
    ref = b.results()
    side = torch.cuda.Stream()
    for _ in range(3):
        b.run(ctypes.c_void_p(side.cuda_stream))
        got = b.results()  # NULL stream = the context's stream
        for k in ("price", "delta", "vega", "rho"):
            assert np.array_equal(ref[k], got[k])
    b.close()


@pytest.mark.parametrize("prec,tuning", [(mc.FP32, (2048, 16, 1, 0, 0)), (mc.FP64, (2048, 16, 1, 0, 0)),
                                         (mc.FP32, (2048, 16, 0, 2, 0)), (mc.FP32, (2048, 16, 1, 1, 0)),
                                         (mc.FP64, (2048, 16, 0, 1, 0)),
                                         (mc.FP32, (2048, 16, 0, 2, 1)), (mc.FP64, (2048, 16, 0, 1, 1)),
                                         (mc.FP32, (2048, 16, 1, 3, 1)), (mc.FP64, (1024, 0, 1, 1, 1))])
def test_graph_and_persistent_match_grouped_launches(engine, prec, tuning):
    # graph replay (with eager work and synchronisations between replays),
    # the persistent static-schedule kernel and the persistent work-queue
    # kernel give the grouped stream launches' results bit for bit on every
    # run (an item's sums do not depend on which workgroup runs it), and
    # their completion times order the micro-batches
    import torch
    reqs = mc.mixed_requests(64, seed=8)
    m = mc.mode(prec, seed=6)
    ref = mc.Batch(engine, reqs, m, tuning=(2048, 16, 0, 0) if tuning[1] else (1024, 0, 0, 0))
    ref.run()
    want = ref.results()
    ref.close()
    g = mc.Batch(engine, reqs, m, tuning=tuning)
    assert g.tuning() == tuning
    for _ in range(4):
        x = torch.randn(1 << 20, device="cuda")
        torch.cuda.synchronize()
        del x
        g.run()
        got = g.results()
        for k in ("price", "delta", "vega", "rho", "price_se", "delta_se", "vega_se", "rho_se"):
            assert np.array_equal(got[k], want[k]), k
        lat = got["latency_us"]
        assert np.all(lat > 0) and np.all(np.diff(lat[::16]) >= 0)
    g.close()


def test_mixed_batch_statistics(engine):
    # GBM requests within 3 SE of Black-Scholes (Greeks too)
    reqs = mc.mixed_requests(200, seed=6)
    b = mc.Batch(engine, reqs, mc.mode(mc.FP32, seed=4))
    b.run()
    out = b.results()
    b.close()
    zs = []
    for i, q in enumerate(reqs):
        if q.kind != mc.REQ_GBM_EUROPEAN or out["price_se"][i] <= 0:
            continue
        if (out["price"][i] / out["price_se"][i]) ** 2 < 50:
            continue
        zs.append((out["price"][i] - orc.bs_call(q.S0, q.K, q.T, q.r, q.sigma)) / out["price_se"][i])
    zs = np.array(zs)
    assert len(zs) > 60 and np.max(np.abs(zs)) < 5 and abs(np.mean(zs * zs) - 1) < 0.5


def test_batch_errors(engine):
    with pytest.raises(mc.FsmcError):
        mc.Batch(engine, [mc.gbm_request(1, 10, 10, 1, .02, .3)], mc.mode())
    with pytest.raises(mc.FsmcError):
        mc.Batch(engine, [mc.gbm_request(100, 10, 10, 1, .02, .3)], mc.mode(), tuning=(1000, 0))
    with pytest.raises(mc.FsmcError):
        mc.Batch(engine, [mc.gbm_request(100, 10, 10, 1, .02, .3)], mc.mode(), tuning=(1024, 0, 2))
    with pytest.raises(mc.FsmcError):
        mc.Batch(engine, [mc.gbm_request(100, 10, 10, 1, .02, .3)], mc.mode(), tuning=(1024, 0, 0, 17))
    with pytest.raises(mc.FsmcError):  # the work queue runs on the persistent grid only
        mc.Batch(engine, [mc.gbm_request(100, 10, 10, 1, .02, .3)], mc.mode(), tuning=(1024, 0, 0, 0, 1))
    b = mc.Batch(engine, [mc.gbm_request(100, 10, 10, 1, .02, .3)], mc.mode())
    with pytest.raises(mc.FsmcError):  # a micro-batch holds at least one request
        b.autotune(requests_per_launch=0)
    b.close()


def test_serving_batch_4096_requests(engine):
    # The serving batch at its full size (4096 requests, FP32, Greeks): GBM
    # requests against Black-Scholes price, delta, vega and rho in units of
    # their standard errors; Heston requests against the FP64 batch on the
    # same stream; one micro-batch, so every request completes together
    reqs = mc.mixed_requests(4096)
    out = {}
    for prec in (mc.FP32, mc.FP64):
        b = mc.Batch(engine, reqs, mc.mode(prec, seed=7), tuning=(4096, 0, 0, 0))
        b.run()
        out[prec] = b.results()
        b.close()
    r32, r64 = out[mc.FP32], out[mc.FP64]
    kind = _fields(reqs, "kind")
    g = (kind == mc.REQ_GBM_EUROPEAN) & (r32["price_se"] > 0)
    g &= (r32["price"] / np.maximum(r32["price_se"], 1e-300)) ** 2 > 50  # enough paths in the money
    zs = {k: [] for k in ("price", "delta", "vega", "rho")}
    for i in np.flatnonzero(g):
        q = reqs[i]
        ref = [orc.bs_call(q.S0, q.K, q.T, q.r, q.sigma)] + orc.bs_call_greeks(q.S0, q.K, q.T, q.r, q.sigma)
        # rho's pathwise estimator is K T e^{-rT} 1{in the money}: with fewer
        # than 5 paths out of the money its standard error is not a
        # Gaussian scale (0, or rounding when every path is in)
        kt = q.K * q.T * np.exp(-q.r * q.T)
        few_out = q.n_paths * (1.0 - r32["rho"][i] / kt) < 5
        for k, v in zip(("price", "delta", "vega", "rho"), ref):
            if k == "rho" and few_out:
                assert abs(r32[k][i] - v) <= 1e-3 * abs(v), (k, i)
            else:
                zs[k].append((r32[k][i] - v) / r32[k + "_se"][i])
    for k, z in zs.items():
        z = np.array(z)
        assert len(z) > 2000 and np.max(np.abs(z)) < 5.5 and abs(np.mean(z * z) - 1) < 0.15, k
    h = kind == mc.REQ_HESTON_ASIAN
    # FP32 against FP64 on the same stream, in standard errors: prices stay
    # within rounding; the pathwise Greeks carry the full-truncation kink
    # (a path whose variance crosses 0 at a different step in FP32 changes
    # its vega contribution), measured median 5e-5, p99 1.4e-2, worst 0.16
    # of ~1165 requests
    d = np.abs(r32["price"][h] - r64["price"][h]) / r64["price_se"][h]
    assert np.max(d) < 0.01
    for k in ("delta", "vega", "rho"):
        d = np.abs(r32[k][h] - r64[k][h]) / np.maximum(r64[k + "_se"][h], 1e-300)
        assert np.median(d) < 1e-3 and np.percentile(d, 99) < 0.05 and np.max(d) < 0.5, k
    lat = r32["latency_us"]
    assert np.all(lat > 0) and np.ptp(lat) == 0
