"""The allocation-free Heston-Asian entry point (fsmc_heston_asian_noalloc)
against the workspace entry point (fsmc_heston_asian_async) it shares its
kernels with: same inputs, same launch shapes, so the results must be
bitwise equal.  Two calls in a row, the second on a smaller input, show that
what the first call left in the shared device array does not leak into the
second; every call is also checked against the FP64 oracle where it applies."""
import numpy as np
import pytest

import mcengine as mc
import oracle as orc

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    S = rng.uniform(80, 120, n)
    cols = [S, np.full(n, 100.0), rng.uniform(.5, 1.5, n), rng.uniform(.03, .06, n), rng.uniform(1, 3, n),
            rng.uniform(.03, .06, n), rng.uniform(.2, .5, n)]
    return cols


def _run_both(cols, n_paths, n_steps, m, greeks, z=None):
    import torch
    dev = torch.device("cuda:0")
    n = cols[0].size
    prm = [torch.tensor(c, dtype=torch.float64, device=dev) for c in cols]
    z1 = z2 = None
    if z is not None:
        z1, z2 = (torch.tensor(a.ravel(), dtype=torch.float64, device=dev) for a in z)
    outs = []
    for kind in ("async", "noalloc"):
        res = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
        conf = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
        g = torch.full((n * 6,), float("nan"), dtype=torch.float64, device=dev) if greeks else None
        if kind == "async":
            ws = torch.empty(mc.heston_asian_workspace(n, n_paths, n_steps, .03, -.7, m), dtype=torch.uint8,
                             device=dev)
            mc.heston_asian_async(prm, n_paths, n_steps, .03, -.7, m, res, conf, ws, z1=z1, z2=z2, out_greeks=g)
        else:
            torch.cuda.synchronize()
            before = torch.cuda.memory_allocated()
            mc.heston_asian_noalloc(prm, n_paths, n_steps, .03, -.7, m, res, conf, z1=z1, z2=z2, out_greeks=g)
            torch.cuda.synchronize()
            assert torch.cuda.memory_allocated() == before
        torch.cuda.synchronize()
        outs.append([t.cpu().numpy() for t in (res, conf) + ((g,) if greeks else ())])
    for a, b in zip(*outs):
        assert np.all(np.isfinite(b))
        assert np.array_equal(a, b)
    return outs[1]


@pytest.mark.parametrize("precision", [mc.FP64, mc.FP32])
def test_noalloc_equals_async_twice_second_smaller(precision):
    print(f"noalloc: Philox, precision {precision}: 12 options x 20000 paths x 40 steps, then 3 x 3000 x 17",
          flush=True)
    m = mc.mode(precision, seed=77)
    big = _inputs(12, 1)
    res, conf = _run_both(big, 20000, 40, m, greeks=False)
    small = _inputs(3, 2)
    res2, conf2 = _run_both(small, 3000, 17, m, greeks=False)
    if precision == mc.FP64:
        ores, oconf = orc.heston_philox(*small, .03, -.7, 3000, 17, seed=77)
        assert np.allclose(res2, ores, rtol=1e-12, atol=0) and np.allclose(conf2, oconf, rtol=1e-10, atol=0)


def test_noalloc_greeks_twice_second_smaller():
    print("noalloc: adjoint Greeks FP64, 4 x 8192 x 36 then 2 x 1024 x 20", flush=True)
    m = mc.mode(mc.FP64, seed=5)
    _run_both(_inputs(4, 3), 8192, 36, m, greeks=True)
    small = _inputs(2, 4)
    res, conf, g = _run_both(small, 1024, 20, m, greeks=True)
    ores, oconf, og = orc.heston_philox(*small, .03, -.7, 1024, 20, seed=5, greeks=True)
    assert np.allclose(res, ores, rtol=1e-12, atol=0)
    assert np.allclose(g.reshape(2, 6)[:, :3], og[:, :3], rtol=1e-9, atol=0)


def test_noalloc_buffer_running_val2_twice_second_smaller():
    # the reference's estimator (chunk sums in the shared array) on caller normals
    print("noalloc: buffer mode + running val2, 3 x 256 x 12 then 2 x 128 x 7", flush=True)
    m = mc.mode(mc.FP64, mc.RNG_BUFFER, quirks=mc.QUIRK_ASIAN_RUNNING_VAL2)
    rng = np.random.default_rng(8)
    for n, paths, steps, seed in ((3, 256, 12, 5), (2, 128, 7, 6)):
        cols = _inputs(n, seed)
        z = (rng.standard_normal((n, paths // 64, steps, 64)), rng.standard_normal((n, paths // 64, steps, 64)))
        res, conf = _run_both(cols, paths, steps, m, greeks=False, z=z)
        ores, oconf = orc.heston_buffer(*cols, .03, -.7, paths, steps, z[0], z[1], quirks=orc.Q_RUNNING_VAL2)
        assert np.allclose(res, ores, rtol=1e-12, atol=0) and np.allclose(conf, oconf, rtol=1e-12, atol=0)


def test_noalloc_rejects_oversized_pricing():
    # 64 x 131072 x 365 with Greeks needs 1.5 GB of checkpoints: refused, nothing launched
    import torch
    dev = torch.device("cuda:0")
    prm = [torch.tensor(c, dtype=torch.float64, device=dev) for c in _inputs(64, 9)]
    res = torch.zeros(64, dtype=torch.float64, device=dev)
    g = torch.zeros(64 * 6, dtype=torch.float64, device=dev)
    with pytest.raises(mc.FsmcError) as e:
        mc.heston_asian_noalloc(prm, 131072, 365, .03, -.7, mc.mode(mc.FP32, seed=1), res, res.clone(), out_greeks=g)
    assert e.value.status == -4
