"""Pins of the CPU oracle (no GPU).

The reference ships no golden vectors or tests, so each oracle function is
pinned here by what the mathematics fixes: published known-answer vectors,
textbook values, closed forms, exact martingale identities and conditional
closed-form quadrature -- each chosen so that a dropped term, a wrong sign or
a swapped operand in the oracle fails at least one check.
"""
import math
import os

import numpy as np
import pytest
from numpy.polynomial.hermite_e import hermegauss
from scipy.stats import norm

import oracle as orc


# ---- Philox4x32-10: Random123 known-answer vectors (kat_vectors) --------------
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_kat(ctr, key, expect):
    assert orc.philox4x32_10(ctr, key) == expect


def test_box_muller_identities():
    # u2 = 0 -> angle 0: z0 = r, z1 = 0; a = 0 -> u1 = fl(1) 2^-32, the
    # smallest uniform; a = 2^32 - 1 rounds up to u1 = 1 -> r = 0.
    z0, z1 = orc.box_muller(0, 0)
    assert z1 == 0.0 and z0 == pytest.approx(math.sqrt(-2 * math.log(2.0 ** -32)), rel=1e-15)
    z0, z1 = orc.box_muller(0xFFFFFFFF, 0x40000000)  # u1 = 1, u2 = 1/4
    assert z0 == 0.0 and z1 == 0.0
    # u2 = 1/4 -> (cos, sin) = (0, 1): all radius in z1.  0x7FFFFFFF rounds
    # to 2^31 in FP32 (24-bit significand), so u1 = 1/2.
    z0, z1 = orc.box_muller(0x7FFFFFFF, 0x40000000)
    r = math.sqrt(2 * math.log(2.0))
    assert abs(z0) < 1e-15 * r and z1 == pytest.approx(r, rel=1e-15)
    # below 2^24 the conversion is exact: u1 = (a | 1) 2^-32
    z0, _ = orc.box_muller(0x00ABCDEE, 0)
    assert z0 == pytest.approx(math.sqrt(-2 * math.log(0x00ABCDEF * 2.0 ** -32)), rel=1e-15)
    # above it, round to nearest even on 24 significant bits: 2^24 + 1 is a
    # tie between 2^24 and 2^24 + 2 and goes to the even significand 2^24
    z0, _ = orc.box_muller(2 ** 24, 0)
    assert z0 == pytest.approx(math.sqrt(-2 * math.log(2.0 ** -8)), rel=1e-15)
    # u2 = fl(b) 2^-32: b = 2^32 - 1 rounds to one full revolution
    z0, z1 = orc.box_muller(0, 0xFFFFFFFF)
    assert z0 == pytest.approx(math.sqrt(64 * math.log(2.0)), rel=1e-15) and abs(z1) < 1e-14


def test_normals_moments():
    z = np.array([orc.normals4(7, 0, 2, p, s, 0) for p in range(2000) for s in range(8)]).ravel()
    n = z.size
    assert abs(z.mean()) < 4 / math.sqrt(n)
    assert abs(z.var() - 1) < 4 * math.sqrt(2 / n)
    assert abs(np.mean(z ** 3)) < 4 * math.sqrt(15 / n)
    assert abs(np.mean(z ** 4) - 3) < 4 * math.sqrt(96 / n)
    # different tags / streams are different draws
    assert orc.normals4(7, 0, 1, 0, 0, 0) != orc.normals4(7, 0, 2, 0, 0, 0)
    assert orc.normals4(7, 0, 1, 0, 0, 0) != orc.normals4(7, 1, 1, 0, 0, 0)


def test_mt19937_pins():
    # MKL's first outputs for vslNewStream(VSL_BRNG_MT19937, 42)
    assert list(orc.mt19937_bits(42, 0, 3)) == [2746317213, 478163327, 107420369]
    # numpy's legacy RandomState([seed]) seeds MT19937 by the same
    # init_by_array; its full-range 32-bit draws are the raw outputs
    for seed in (43, 7, 2**32 - 1):
        want = np.random.RandomState([seed]).randint(0, 2**32, size=2000, dtype=np.uint32)
        assert np.array_equal(orc.mt19937_bits(seed, 0, 2000), want)
        assert np.array_equal(orc.mt19937_bits(seed, 1247, 700), want[1247:1947])
    z = orc.mt19937_gaussian(42, 0, 200000)
    n = z.size
    assert abs(z.mean()) < 4 / math.sqrt(n) and abs(z.var() - 1) < 4 * math.sqrt(2 / n)
    assert abs(np.mean(z ** 4) - 3) < 4 * math.sqrt(96 / n)
    # pairs: z0 and z1 of a pair share the radius sqrt(-2 ln u1)
    b = orc.mt19937_bits(42, 0, 8).astype(np.float64) * 2.0 ** -32
    r2 = z[:8:2] ** 2 + z[1:8:2] ** 2
    assert np.allclose(r2, -2 * np.log(b[::2]), rtol=1e-13)


# ---- Black-Scholes and lattices --------------------------------------------------
def test_bs_textbook():
    # Hull, Options Futures and Other Derivatives, example 15.6: c = 4.76.
    assert orc.bs_call(42, 40, 0.5, 0.1, 0.2) == pytest.approx(4.7594, abs=1e-4)
    assert orc.bs_call(100, 100, 1, 0.05, 0.2) == pytest.approx(10.450583572185565, rel=1e-13)
    d1 = (math.log(42 / 40) + (0.1 + 0.02) * 0.5) / (0.2 * math.sqrt(0.5))
    g = orc.bs_call_greeks(42, 40, 0.5, 0.1, 0.2)
    assert g[0] == pytest.approx(norm.cdf(d1), rel=1e-13)
    # Greeks against central differences of the price
    h = 1e-5
    fd = [(orc.bs_call(42 + h, 40, .5, .1, .2) - orc.bs_call(42 - h, 40, .5, .1, .2)) / (2 * h),
          (orc.bs_call(42, 40, .5, .1, .2 + h) - orc.bs_call(42, 40, .5, .1, .2 - h)) / (2 * h),
          (orc.bs_call(42, 40, .5, .1 + h, .2) - orc.bs_call(42, 40, .5, .1 - h, .2)) / (2 * h)]
    assert np.allclose(g, fd, rtol=1e-7)


def test_crr_european_converges_to_bs():
    for S, X, T, r, s in [(42, 40, .5, .1, .2), (20, 80, 10, .02, .3), (30, 1, .25, .02, .3)]:
        bs = orc.bs_call(S, X, T, r, s)
        c2 = orc.crr_european_call(S, X, T, r, s, 2048)
        c3 = orc.crr_european_call(S, X, T, r, s, 2049)
        assert abs(c2 - bs) < 2e-3 * max(bs, 1e-2)
        # CRR error oscillates between odd and even step counts; their mean
        # converges much faster.
        assert abs(0.5 * (c2 + c3) - bs) < 5e-4 * max(bs, 1.0)


def test_crr_american_put():
    # Longstaff & Schwartz (2001) Table 1, finite-difference column.
    assert orc.crr_american_put(36, 40, 1, .06, .2, 4096) == pytest.approx(4.478, abs=1e-2)
    assert orc.crr_american_put(40, 40, 1, .06, .2, 4096) == pytest.approx(2.314, abs=1e-2)
    assert orc.crr_american_put(44, 40, 2, .06, .4, 4096) == pytest.approx(5.647, abs=2e-2)
    # converged to 1e-3 between 4096 and 8192 steps
    a = orc.crr_american_put(36, 40, 1, .06, .2, 4096)
    b = orc.crr_american_put(36, 40, 1, .06, .2, 8192)
    assert abs(a - b) < 1e-3
    # r = 0: early exercise is worthless, American put = European put (parity)
    eu_put = orc.bs_call(36, 40, 1, 0.0, .2) - 36 + 40
    assert orc.crr_american_put(36, 40, 1, 0.0, .2, 4096) == pytest.approx(eu_put, abs=2e-3)
    # American >= European
    eu_put6 = orc.bs_call(36, 40, 1, .06, .2) - 36 + 40 * math.exp(-.06)
    assert orc.crr_american_put(36, 40, 1, .06, .2, 4096) > eu_put6 + 0.5


def test_geometric_asian_closed_form():
    # N = 1 monitoring date is a European call.
    assert orc.geometric_asian_call(100, 95, 1, .03, .25, 1) == pytest.approx(
        orc.bs_call(100, 95, 1, .03, .25), rel=1e-13)


# ---- GBM oracle --------------------------------------------------------------------
def test_gbm_oracle_vs_black_scholes():
    rng = np.random.default_rng(1)
    n = 16
    S = rng.uniform(5, 30, n)
    X = rng.uniform(1, 100, n)
    T = rng.uniform(.25, 10, n)
    res, conf, g = orc.gbm_philox(S, X, T, .02, .3, 1 << 16, seed=11, greeks=True)
    se = conf / 1.96
    for i in range(n):
        bs = orc.bs_call(S[i], X[i], T[i], .02, .3)
        assert abs(res[i] - bs) < 4 * se[i] + 1e-8
        gb = orc.bs_call_greeks(S[i], X[i], T[i], .02, .3)
        for k in range(3):
            assert abs(g[i, k] - gb[k]) < 4.5 * g[i, 3 + k] + 1e-7


def test_gbm_buffer_oracle_matches_philox_oracle():
    # The buffer estimator fed with the Philox normals (converted to log2
    # samples) reproduces the Philox estimator: same arithmetic, same draws.
    S, X, T = np.array([20.0, 7.0]), np.array([18.0, 9.0]), np.array([2.0, .5])
    n_paths, r, sig = 4096, .02, .3
    z = np.array([[orc.normals4(5, 0, 1, q, 0, o)[j] for q in range(n_paths // 4) for j in range(4)]
                  for o in range(2)])
    L2E = 1.4426950408889634074
    mu = (L2E * (r - .5 * sig * sig) * T)[:, None]
    vol = (L2E * sig * np.sqrt(T))[:, None]
    samples = (mu + vol * z).astype(np.float32)
    rb, cb = orc.gbm_buffer(S, X, T, r, sig, samples)
    rp, cp = orc.gbm_philox(S, X, T, r, sig, n_paths, seed=5)
    # only the float rounding of the samples differs
    assert np.allclose(rb, rp, rtol=2e-6)
    assert np.allclose(cb, cp, rtol=2e-5)


def test_gbm_control_variate_small_strike_is_exact():
    # K far below every S_T: y = (S_T - K) = u + S e^{rT} - K on every path,
    # so beta = 1 and the estimator is S - K e^{-rT} with zero variance.  A
    # wrong sign of beta, an uncentred control or a dropped discount fails.
    S, X, T = np.array([20.0, 7.5]), np.array([1e-3, 2e-3]), np.array([1.0, 3.0])
    r = .02
    res, conf = orc.gbm_philox_cv(S, X, T, r, .3, 8192, seed=3)
    assert np.allclose(res, S - X * np.exp(-r * T), rtol=1e-12, atol=0)
    assert np.all(conf < 1e-9 * S)
    # the buffer variant on float samples
    rng = np.random.default_rng(4)
    L2E = 1.4426950408889634074
    z = rng.standard_normal((2, 4096))
    samples = ((L2E * (r - .045) * T)[:, None] + (L2E * .3 * np.sqrt(T))[:, None] * z).astype(np.float32)
    rb, cb = orc.gbm_buffer_cv(S, X, T, r, samples)
    # the float samples' mean of S_T is not exactly S e^{rT}: the estimator
    # is S - K e^{-rT} plus the sample mean's deviation times (1 - beta) = 0
    assert np.allclose(rb, S - X * np.exp(-r * T), rtol=1e-12, atol=0)


def test_gbm_control_variate_vs_black_scholes():
    rng = np.random.default_rng(9)
    n = 24
    S = rng.uniform(5, 30, n)
    X = rng.uniform(1, 40, n)
    T = rng.uniform(.25, 5, n)
    res, conf = orc.gbm_philox_cv(S, X, T, .02, .3, 1 << 15, seed=21)
    pres, pconf = orc.gbm_philox(S, X, T, .02, .3, 1 << 15, seed=21)
    se = conf / 1.96
    for i in range(n):
        bs = orc.bs_call(S[i], X[i], T[i], .02, .3)
        # deep in the money every path may finish above K: the estimator is
        # then exactly S - K e^{-rT} (zero variance) and misses the put value
        # bs - (S - K e^{-rT}) that no path saw
        put = bs - (S[i] - X[i] * np.exp(-.02 * T[i]))
        assert abs(res[i] - bs) < 4 * se[i] + 1e-8 + (put if put < 1e-4 else 0.0)
    # never worse than the plain estimator on the same paths, much better in the money
    assert np.all(conf <= pconf * (1 + 1e-12))
    itm = S > 1.5 * X
    assert itm.any() and np.all(conf[itm] < 0.5 * pconf[itm])


# ---- Heston oracle -------------------------------------------------------------------
HP = dict(S=[100.0], X=[100.0], T=[1.0], v0=[0.09], kappa=[2.0], theta=[0.04], xi=[1.0])


def _heston_n2_quadrature(S0, K, T, r, v0, kappa, theta, xi, rho, var_as_vol=False, n=160):
    """E[e^{-rT} max((S1+S2)/2 - K, 0)] for two Euler steps: the inner
    expectation over the second spot normal is a Black-Scholes-type closed
    form; the outer one over (z1_0, z2_0) is 2-D Gauss-Hermite."""
    x, w = hermegauss(n)
    w = w / math.sqrt(2 * math.pi)
    a0, b0 = np.meshgrid(x, x, indexing="ij")
    W = np.outer(w, w)
    dt = T / 2
    s0 = v0 if var_as_vol else math.sqrt(v0)
    S1 = S0 * np.exp((r - .5 * s0 * s0) * dt + s0 * math.sqrt(dt) * a0)
    zc = rho * a0 + math.sqrt(1 - rho * rho) * b0
    v1 = v0 + kappa * dt * (theta - v0) + xi * math.sqrt(v0 * dt) * zc
    v1p = np.maximum(v1, 0.0)
    s1 = v1p if var_as_vol else np.sqrt(v1p)
    m = (r - .5 * s1 * s1) * dt
    s = s1 * math.sqrt(dt)
    Kp = 2 * K - S1
    fwd = S1 * np.exp(r * dt)
    with np.errstate(divide="ignore", invalid="ignore"):
        d2 = (np.log(S1 / np.where(Kp > 0, Kp, 1.0)) + m) / np.where(s > 0, s, 1.0)
        bs = fwd * norm.cdf(d2 + s) - Kp * norm.cdf(d2)
    inner = np.where(Kp <= 0, fwd - Kp, np.where(s > 0, bs, np.maximum(fwd - Kp, 0.0)))
    return math.exp(-r * T) * 0.5 * float(np.sum(W * inner))


@pytest.mark.parametrize("var_as_vol", [False, True])
def test_heston_oracle_two_step_quadrature(var_as_vol):
    r, rho = 0.03, -0.7
    q = orc.Q_VAR_AS_VOL if var_as_vol else 0
    res, conf = orc.heston_philox(r=r, rho=rho, n_paths=1 << 22, n_steps=2, seed=3, quirks=q, **HP)
    se = conf[0] / 1.96
    exact = _heston_n2_quadrature(100, 100, 1, r, .09, 2, .04, 1.0, rho, var_as_vol)
    assert abs(res[0] - exact) < 4 * se
    # the pin discriminates the correlation sign and the variance coefficients:
    # each alternative lies more than 8 SE away, so it would fail the 4-SE check
    for alt in (dict(rho=+0.7), dict(kappa=0.0), dict(xi=0.0)):
        kw = dict(S0=100, K=100, T=1, r=r, v0=.09, kappa=2, theta=.04, xi=1.0, rho=rho)
        kw.update(alt)
        assert abs(_heston_n2_quadrature(var_as_vol=var_as_vol, **kw) - exact) > 8 * se


def test_heston_oracle_martingale_mean():
    # K = 0: the call pays the average, whose expectation is exact:
    # e^{-rT} S0 (1/N) sum_k e^{r k dt} (the log-Euler step is a martingale).
    N, r = 12, 0.05
    for q in (0, orc.Q_VAR_AS_VOL):
        res, conf = orc.heston_philox([100.0], [0.0], [2.0], [.04], [2.0], [.04], [.3], r, -.7, 1 << 17, N,
                                      seed=9, quirks=q)
        dt = 2.0 / N
        exact = math.exp(-r * 2) * 100 * np.mean(np.exp(r * dt * np.arange(1, N + 1)))
        assert abs(res[0] - exact) < 4 * conf[0] / 1.96


def test_heston_oracle_one_step_is_black_scholes():
    # one monitoring date and xi = 0: a Black-Scholes call with sigma = sqrt(v0)
    res, conf = orc.heston_philox([100.0], [95.0], [1.0], [.0625], [3.0], [.5], [0.0], .03, -.7, 1 << 20, 1,
                                  seed=4)
    assert abs(res[0] - orc.bs_call(100, 95, 1, .03, .25)) < 4 * conf[0] / 1.96


def test_heston_buffer_oracle_matches_philox_oracle():
    # Buffer (reference-layout) estimator fed with the Philox draws equals the
    # Philox estimator to rounding: checks the [opt][chunk][step][lane] layout.
    n_paths, n_steps = 128, 6
    P = dict(S=[100.0, 90.0], X=[100.0, 95.0], T=[1.0, .5], v0=[.04, .05], kappa=[2.0, 1.0],
             theta=[.04, .06], xi=[.3, .5])
    z1 = np.zeros((2, n_paths // 64, n_steps, 64))
    z2 = np.zeros_like(z1)
    for o in range(2):
        for p in range(n_paths):
            for k in range(0, n_steps, 2):
                n = orc.normals4(21, 0, 2, p, k // 2, o)
                z1[o, p // 64, k, p % 64], z2[o, p // 64, k, p % 64] = n[0], n[1]
                z1[o, p // 64, k + 1, p % 64], z2[o, p // 64, k + 1, p % 64] = n[2], n[3]
    rb, cb = orc.heston_buffer(r=.03, rho=-.7, n_paths=n_paths, n_steps=n_steps, z1=z1, z2=z2, **P)
    rp, cp = orc.heston_philox(r=.03, rho=-.7, n_paths=n_paths, n_steps=n_steps, seed=21, **P)
    assert np.allclose(rb, rp, rtol=1e-12) and np.allclose(cb, cp, rtol=1e-10)


def test_heston_oracle_greeks_vs_fd():
    # Tangent-mode pathwise Greeks against central differences of the same
    # common-random-number estimator.
    n_paths, N, r = 1 << 14, 24, 0.03
    P = dict(S=[100.0], X=[100.0], T=[1.0], v0=[.04], kappa=[2.0], theta=[.04], xi=[.3])
    res, conf, g = orc.heston_philox(r=r, rho=-.7, n_paths=n_paths, n_steps=N, seed=5, greeks=True, **P)

    def price(**kw):
        q = dict(P, r=r)
        q.update(kw)
        return orc.heston_philox(rho=-.7, n_paths=n_paths, n_steps=N, seed=5, **q)[0][0]

    h = 1e-4
    fd_delta = (price(S=[100 + h]) - price(S=[100 - h])) / (2 * h)
    s0 = 0.2
    fd_vega = (price(v0=[(s0 + h) ** 2]) - price(v0=[(s0 - h) ** 2])) / (2 * h)
    fd_rho = (price(r=r + h) - price(r=r - h)) / (2 * h)
    assert g[0, 0] == pytest.approx(fd_delta, rel=2e-3)
    assert g[0, 1] == pytest.approx(fd_vega, rel=2e-3)
    assert g[0, 2] == pytest.approx(fd_rho, rel=2e-3)


# ---- LIBOR oracle ----------------------------------------------------------------------
def _libor_fd(z, L0, h=1e-6):
    fd = np.zeros(80)
    for i in range(80):
        a, b = L0.copy(), L0.copy()
        a[i] += h
        b[i] -= h
        fd[i] = (orc.libor_path(z, a)[0] - orc.libor_path(z, b)[0]) / (2 * h)
    return fd


def test_libor_oracle_adjoint_matches_finite_differences():
    # path_calc_b2 + portfolio_b with a separate L_b is the exact derivative
    # of the per-path value: central differences on each initial rate.
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(6):
        z = rng.standard_normal(40)
        L0 = rng.uniform(0.04, 0.06, 80)
        v, Lb = orc.libor_path(z, L0)
        if v == 0.0:
            continue
        fd = _libor_fd(z, L0)
        assert np.max(np.abs(Lb - fd)) < 1e-7 * np.max(np.abs(fd))
        checked += 1
    assert checked >= 3


def test_libor_oracle_inplace_quirk_reproduces_reference_defect():
    # liborSwaptionGreeks.cpp:178 overwrites L[n] with L_b[n] and :181 reads
    # it back: deltas 1..78 go wrong, 0 and 79 (the one the sample prints) stay
    # exact, and the value is unaffected.
    rng = np.random.default_rng(3)
    for _ in range(6):
        z = rng.standard_normal(40)
        L0 = np.full(80, 0.051)
        v, Lb = orc.libor_path(z, L0)
        vq, Lq = orc.libor_path(z, L0, quirks=orc.Q_LIBOR_INPLACE)
        assert vq == v
        if v == 0.0:
            continue
        fd = _libor_fd(z, L0)
        scale = np.max(np.abs(fd))
        assert abs(Lq[0] - fd[0]) < 1e-7 * scale and abs(Lq[79] - fd[79]) < 1e-7 * scale
        assert np.max(np.abs(Lq[1:79] - fd[1:79])) > 0.1 * scale
        return
    raise AssertionError("no in-the-money path drawn")


def test_libor_oracle_zero_vol_closed_form():
    # lambda = 0: rates stay on the flat curve, the payer swap value is
    # B_m + K S_m - 1 = S_m (K - L0) with B_m = (1 + d L0)^-(m+1) and
    # S_m = d sum_{k<=m} B_k; swaptions pay 100 max(-(.), 0) discounted by
    # (1 + d L0)^-40.
    for L0 in (0.05, 0.051, 0.047):
        d = 0.25
        B = (1 + d * L0) ** -np.arange(1, 41)
        S = d * np.cumsum(B)
        mats = [4, 4, 4, 8, 8, 8, 20, 20, 20, 28, 28, 28, 40, 40, 40]
        rates = np.array([.045, .05, .055] * 5, dtype=np.float32).astype(np.float64)
        val = sum(100 * max(-(B[m - 1] + K * S[m - 1] - 1), 0.0) for m, K in zip(mats, rates))
        val *= (1 + d * L0) ** -40
        v, _ = orc.libor_path(np.random.default_rng(0).standard_normal(40), L0, lam=0.0)
        assert v == pytest.approx(val, rel=1e-13)


def test_libor_oracle_statistics_stable():
    # Philox mode: the value is within 4 SE across independent streams
    a = orc.libor(4096, L0=0.05, seed=1)
    b = orc.libor(4096, L0=0.05, seed=2)
    assert abs(a[0] - b[0]) < 4 * math.hypot(a[1], b[1])


# ---- Longstaff-Schwartz oracle ---------------------------------------------------------
def test_chol4_solve():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((6, 4))
    G = A.T @ A
    b = rng.standard_normal(4)
    assert np.allclose(orc.chol4_solve(G, b), np.linalg.solve(G, b), rtol=1e-12)
    assert orc.chol4_solve(np.diag([1.0, 1.0, 0.0, 1.0]), b) is None


def test_lsmc_oracle_longstaff_schwartz_table():
    # Longstaff & Schwartz (2001) Table 1 (50 exercise dates per year, 100k
    # paths): LSM 4.472 / 2.313 / 1.110 for S0 = 36 / 40 / 44, K = 40, r = .06,
    # sigma = .2, T = 1.  The oracle's prices sit within 3 SE + the regression
    # bias of the lattice value (LSMC is biased low).
    S0 = [36.0, 40.0, 44.0]
    n = len(S0)
    p, se, eu = orc.lsmc(S0, [40.0] * n, [.2] * n, [.06] * n, [1.0] * n, [50] * n, [100000] * n, seed=1)
    for i in range(n):
        crr = orc.crr_american_put(S0[i], 40, 1, .06, .2, 4096)
        assert crr - 3 * se[i] - 0.02 < p[i] < crr + 3 * se[i]
        # European put on the same paths vs Black-Scholes parity
        bs_put = orc.bs_call(S0[i], 40, 1, .06, .2) - S0[i] + 40 * math.exp(-.06)
        assert abs(eu[i] - bs_put) < 0.05
    assert p == pytest.approx([4.472, 2.313, 1.110], abs=0.03)


def test_lsmc_oracle_zero_rate_no_early_exercise():
    # r = 0, no dividends: early exercise of a put is never optimal; LSMC
    # can only lose by exercising, so American <= European + noise, and it
    # stays close to the European value.
    p, se, eu = orc.lsmc([100.0], [100.0], [.3], [0.0], [1.0], [50], [50000], seed=3)
    assert p[0] <= eu[0] + 3 * se[0]
    assert p[0] > eu[0] - 0.05


def test_lsmc_ref_oracle_reference_quirks():
    # The sample's algorithm: no rate, so its "American" and "European" prices
    # estimate the same quantity; the European leg is a plain GBM put average
    # equal to Black-Scholes with r = 0 within MC error.
    rng = np.random.default_rng(5)
    n_paths, n_steps = 16384, 64
    z = rng.standard_normal(n_paths * n_steps)
    am, eu = orc.lsmc_ref(n_paths, n_steps, 100.0, 110.0, .25, 2.0, z)
    bs_put = orc.bs_call(100, 110, 2, 0.0, .25) - 100 + 110
    assert abs(eu - bs_put) < 0.5
    assert abs(am - eu) < 1.5
    # a path that never enters the money has zero cash flow; with strike 0 the
    # put is worthless in both legs
    am0, eu0 = orc.lsmc_ref(256, 8, 100.0, 1e-9, .25, 2.0, z[:256 * 8])
    assert am0 == 0.0 and eu0 == 0.0


def test_crr_bermudan_put():
    # every lattice step an exercise date -> the American lattice
    am = orc.crr_american_put(36, 40, 1, .06, .2, 1024)
    assert orc.crr_bermudan_put(36, 40, 1, .06, .2, 1024, 1) == pytest.approx(am, rel=1e-14)
    # one date (t = 0 and T) -> max(K - S0, European)
    eu = orc.crr_european_call(36, 40, 1, .06, .2, 2048) - 36 + 40 * math.exp(-.06)
    assert orc.crr_bermudan_put(36, 40, 1, .06, .2, 1, 2048) == pytest.approx(max(4.0, eu), abs=2e-3)
    # more exercise dates are worth more, bounded by the American
    b10 = orc.crr_bermudan_put(36, 40, 1, .06, .2, 10, 400)
    b50 = orc.crr_bermudan_put(36, 40, 1, .06, .2, 50, 80)
    assert eu < b10 < b50 < orc.crr_american_put(36, 40, 1, .06, .2, 4000) + 1e-3


# ---- MKL-compatible MRG32k3a + inverse CDF (the generateRand seam) -----------
# Printed by MKL 2021.4 for vslNewStream(VSL_BRNG_MRG32K3A, 12345): the first
# uniforms (vdRngUniform) and the first normals of vdRngGaussian with
# VSL_RNG_METHOD_GAUSSIAN_ICDF, i.e. z[0..5] of LiborMarketModel's buffer.
MKL_MRG_U = [0.67025736022828808, 0.55588071611222578, 0.21607400713476063]
MKL_MRG_Z = [0.44062392580367954, 0.14053339627998879, -0.78552125308716181,
             2.1380353354757777, -0.77725596281159304, 0.090591905930727717]


def test_mrg32k3a_reproduces_mkl_uniforms_and_normals():
    u = orc.mrg32k3a_uniform(12345, 0, 3)
    assert list(u) == MKL_MRG_U  # bit for bit
    z = orc.mrg32k3a_gaussian(12345, 0, 6)
    # MKL's ICDF itself is off the exact inverse by up to ~7.3e-9 relative
    assert np.max(np.abs(z - MKL_MRG_Z) / np.abs(MKL_MRG_Z)) < 3e-9
    # vslSkipAheadStream(4) continues with normals 4 and 5
    assert np.array_equal(orc.mrg32k3a_gaussian(12345, 4, 2), z[4:6])


def _mrg_bigint(seed, n):
    m1, m2 = 4294967087, 4294944443
    x, y, out = [seed % m1, 1, 1], [1, 1, 1], []
    for _ in range(n):
        xn = (1403580 * x[1] - 810728 * x[0]) % m1
        yn = (527612 * y[2] - 1370589 * y[0]) % m2
        x, y = [x[1], x[2], xn], [y[1], y[2], yn]
        z = (xn - yn) % m1
        out.append(z * (1.0 / m1) if z else 0.5 / m1)
    return np.array(out)


def test_mrg32k3a_recurrence_and_skip_ahead():
    ref = _mrg_bigint(987654321, 60000)  # exact integer arithmetic
    assert np.array_equal(orc.mrg32k3a_uniform(987654321, 0, 60000), ref)
    for off in (1, 3, 1000, 41237, 59000):  # matrix-power skip == stepping
        assert np.array_equal(orc.mrg32k3a_uniform(987654321, off, 1000), ref[off:off + 1000])


def test_icdf_matches_exact_inverse():
    from scipy.special import ndtri
    u = np.concatenate([np.logspace(-300, -1, 300), np.linspace(0.01, 0.99, 999), 1 - np.logspace(-16, -2, 100)])
    x = np.array([orc.icdf(v) for v in u])
    assert np.max(np.abs(x - ndtri(u)) / np.maximum(np.abs(ndtri(u)), 1e-300)) < 2e-15
    # the three branches meet continuously; odd symmetry about 1/2
    for v in (0.075, 0.5 - 0.425, np.exp(-25.0)):
        assert orc.icdf(v) == pytest.approx(ndtri(v), rel=1e-15)
    assert orc.icdf(0.3) == -orc.icdf(0.7) or abs(orc.icdf(0.3) + orc.icdf(0.7)) < 1e-15


def test_lsmc_oracle_bridge_paths_have_gbm_law():
    # The batched LSMC builds paths backwards by the Brownian bridge: its
    # European leg must be the Black-Scholes put, for short and long maturities
    S0, K, sig, r = np.array([90.0, 100.0, 110.0]), np.full(3, 100.0), np.array([.2, .3, .4]), np.full(3, .03)
    T, steps, paths = np.array([.5, 1.0, 2.0]), np.array([17, 64, 9]), np.full(3, 200000)
    pr, se, eu = orc.lsmc(S0, K, sig, r, T, steps, paths, seed=5)
    bs_put = np.array([orc.bs_call(*a) for a in zip(S0, K, T, r, sig)]) - S0 + K * np.exp(-r * T)
    sd = np.array([np.sqrt(np.var(np.maximum(K[i] - S0[i] * np.exp((r[i] - sig[i] ** 2 / 2) * T[i] + sig[i]
                                                                        * np.sqrt(T[i]) * np.random.default_rng(i)
                                                                        .standard_normal(200000)), 0))
                           * np.exp(-2 * r[i] * T[i]) / 200000) for i in range(3)])
    assert np.all(np.abs(eu - bs_put) < 4 * sd)
    assert np.all(pr >= eu - 1e-12)


# ---- MKL stream fixtures (tests/golden/mkl_streams.npz, written by
# tests/golden/make_mkl_fixtures.py from Intel oneAPI MKL 2021.4) -------------
def _mkl():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mkl_streams.npz")
    return np.load(path)


def test_mrg32k3a_against_mkl_fixture():
    d = _mkl()
    assert np.array_equal(orc.mrg32k3a_uniform(12345, 0, 2048), d["mrg12345_uniform"])
    f = d["mrg12345_icdf"]
    assert np.max(np.abs(orc.mrg32k3a_gaussian(12345, 0, 2048) - f) / np.abs(f)) < 8e-9
    f = d["mrg12345_skip1000003_icdf"]
    assert np.max(np.abs(orc.mrg32k3a_gaussian(12345, 1000003, 256) - f) / np.abs(f)) < 8e-9


def test_mt19937_against_mkl_fixture():
    # 2048 outputs: three twists of the 624-word state
    d = _mkl()
    for seed in (42, 43):
        assert np.array_equal(orc.mt19937_bits(seed, 0, 2048).astype(np.uint32), d[f"mt{seed}_bits"])
        assert np.max(np.abs(orc.mt19937_gaussian(seed, 0, 2048) - d[f"mt{seed}_boxmuller2"])) < 3e-8


def test_sfmt19937_against_mkl_fixture():
    # MKL's SFMT19937 is init_by_array({seed}) SFMT; its ICDF reads each
    # output as a signed integer (u = int32 2^-32 + 1/2): read unsigned, the
    # normals would be off by O(1) for half the draws
    from scipy.special import ndtri
    d = _mkl()
    assert np.array_equal(orc.sfmt19937_bits(777, 2048), d["sfmt777_bits"])
    z = ndtri(orc.sfmt19937_icdf_uniforms(777, 2048))
    assert np.max(np.abs(z - d["sfmt777_icdf_float"])) < 2e-5


# ---- Pins of the reference-algorithm quirks -----------------------------------------
# Each pin_* function raises AssertionError when the oracle disagrees with what
# the reference's code fixes; tests/test_oracle_mutants.py plants the named
# mistakes in copies of the oracle and shows that each pin rejects them.
LSMC_K, LSMC_SPOT, LSMC_VOL, LSMC_T = 110.0, 100.0, .25, 2.0


def _lsmc_z_from_spots(S):
    """Normals that drive the reference's log2-domain path recursion
    (utilities.cpp:70-83) through the given spots S[date][path], date 0 = spot."""
    n = S.shape[0] - 1
    dt = LSMC_T / n
    return ((np.log(S[1:] / S[:-1]) + .5 * LSMC_VOL ** 2 * dt) / (LSMC_VOL * math.sqrt(dt))).ravel()


def _ridge_fit(x, cf):
    """The sample's regression (main.cpp:236-261): normal equations of the
    quadratic basis over the given in-the-money points, 0.01 added to each of
    the five moments, solved here by LU instead of the cofactor inverse."""
    B = np.stack([np.ones_like(x), x, x * x], axis=1)
    return np.linalg.solve(B.T @ B + 0.01, B.T @ cf)


def pin_lsmc_ref_two_dates():
    """numSteps = 2: the backward loop (main.cpp:131) visits only s = 1 =
    numSteps - 1, where every cash flow is the terminal payoff max(K - S_1, 0)
    scaled by 1/numPaths (:195-199); the final average (:274-278) divides by
    numPaths again.  So American = sum_p max(K - S_1p, 0) / N^2 exactly and
    European = sum_p max(K - S_2p, 0) / N.  A dropped 1/N is off by a factor N,
    a terminal date taken at numSteps uses S_2 (and a second regression)."""
    n, N = 2, 4096
    z = np.random.default_rng(17).standard_normal(n * N)
    am, eu, beta = orc.lsmc_ref(N, n, LSMC_SPOT, LSMC_K, LSMC_VOL, LSMC_T, z, betas=True)
    dt = LSMC_T / n
    S1 = LSMC_SPOT * np.exp(-.5 * LSMC_VOL ** 2 * dt + LSMC_VOL * math.sqrt(dt) * z[:N])
    S2 = S1 * np.exp(-.5 * LSMC_VOL ** 2 * dt + LSMC_VOL * math.sqrt(dt) * z[N:])
    assert am == pytest.approx(np.sum(np.maximum(LSMC_K - S1, 0)) / N ** 2, rel=1e-12)
    assert eu == pytest.approx(np.sum(np.maximum(LSMC_K - S2, 0)) / N, rel=1e-12)
    itm = S1 <= LSMC_K
    want = _ridge_fit(S1[itm], np.maximum(LSMC_K - S1[itm], 0) / N)
    assert np.allclose(beta[1], want, rtol=1e-6, atol=0)


# Five hand-placed paths over four dates (rows: date 1..4), strike 110.
LSMC_CASE = np.array([[108., 88., 106., 116., 119.],
                      [85., 111., 86., 103., 119.],
                      [103., 80., 81., 87., 97.],
                      [89., 128., 123., 85., 107.]])


def pin_lsmc_ref_lag_and_ridge():
    """numSteps = 4, the five paths of LSMC_CASE, strike 110.

    Date 3 (the sweep's terminal date): every path is in the money, cash flows
    (110 - S_3)/5 = 1.4, 6, 5.8, 4.6, 2.6; their ridged fit beta(3) is what
    date 2 decides with.  Date 2: paths 0, 2, 3 are in the money (85, 86,
    103); beta(3) gives them continuation values 5.1, 4.9, 1.3 against
    payoffs 25, 24, 7, so all three exercise.  Their cash flows 110 - x are
    exactly linear in the spot, so without the ridge the date-2 fit would be
    (110, -1, 0) exactly and every later decision a rounding tie; the ridge
    (main.cpp:244-246) on this ill-conditioned 3-point Gram matrix (condition
    2e10) moves it to (-1.349, 1.384, -0.0127).  Date 1 decides with that
    lagged fit (main.cpp:176-190 use the betas of :259-261 from the previous
    iteration): for the in-the-money paths 0 (108), 1 (88), 2 (106) the fit
    minus the payoff is -1.43, +0.48, -0.79, so paths 0 and 2 exercise
    (payoffs 2 and 4) and path 1 (out of the money at date 2) keeps its
    terminal cash flow 6.
    American = (2 + 6 + 4 + 7 + 2.6)/5 = 4.32; European = (21 + 0 + 0 + 25 +
    3)/5 = 9.8.  Planted mistakes move it: the date-1 decision with a fit of
    date 1 itself gives 16.12, no 1/N on the terminal flows 15.6, the terminal
    at date 4 15.6, and no ridge leaves beta(2) at (110, -1, 0)."""
    N, n = LSMC_CASE.shape[1], LSMC_CASE.shape[0]
    S = np.vstack([np.full(N, LSMC_SPOT), LSMC_CASE])
    am, eu, beta = orc.lsmc_ref(N, n, LSMC_SPOT, LSMC_K, LSMC_VOL, LSMC_T, _lsmc_z_from_spots(S), betas=True)
    x3 = LSMC_CASE[2]
    b3 = _ridge_fit(x3, (LSMC_K - x3) / N)
    assert np.allclose(beta[3], b3, rtol=1e-6, atol=0)
    x2 = LSMC_CASE[1][[0, 2, 3]]
    cont = b3[0] + b3[1] * x2 + b3[2] * x2 ** 2
    assert np.all(LSMC_K - x2 > cont + 5)  # all three exercise at date 2
    b2 = _ridge_fit(x2, LSMC_K - x2)
    assert np.allclose(beta[2], b2, rtol=1e-4, atol=0)
    assert abs(b2[0] - 110) > 50  # the ridge matters here
    x1 = LSMC_CASE[0][:3]
    gap = b2[0] + b2[1] * x1 + b2[2] * x1 ** 2 - (LSMC_K - x1)
    assert gap[0] < -1 and gap[1] > 0.4 and gap[2] < -0.7
    assert am == pytest.approx((2 + 6 + 4 + 7 + 2.6) / 5, rel=1e-10)
    assert eu == pytest.approx((21 + 25 + 3) / 5, rel=1e-12)


def _libor_linear_value(L0):
    """Every payer swaption exercised: the spot-measure value is linear in
    bond prices, (P_40 - P_{41+m} - K d sum_k P_{41+k}) per swaption, with
    P_j = (1 + d L0)^-j on the flat initial curve."""
    d = 0.25
    B = (1 + d * L0) ** -np.arange(1, 41)
    S = d * np.cumsum(B)
    mats = [4, 4, 4, 8, 8, 8, 20, 20, 20, 28, 28, 28, 40, 40, 40]
    rates = np.array([.045, .05, .055] * 5, dtype=np.float32).astype(np.float64)
    val = sum(100 * max(-(B[m - 1] + K * S[m - 1] - 1), 0.0) for m, K in zip(mats, rates))
    return val * (1 + d * L0) ** -40


def pin_libor_spot_measure_drift():
    """At L0 = 0.3 every swaption of the book is deep in the money on every
    path, so the payoff is linear in the time-T_40 bond prices and the
    discounted value must equal the lambda = 0 closed form: bonds deflated by
    the spot numeraire are martingales under the spot-measure drift of
    liborSwaptionGreeks.cpp:70-71.  2^17 paths, seed 3, within 3 SE.  The
    log-Euler scheme's own bias is +0.07 % (lambda 0.1) and +0.26 % (lambda
    0.2) of the value, measured at 2^20 paths: about 0.7 and 1.2 SE here.  The
    drift dropped lands at 4.8 and -30 SE, its sign flipped at -21 and -250
    SE, the -con1/2 term's sign flipped at -81 and -172 SE."""
    want = _libor_linear_value(0.3)
    for lam in (0.1, 0.2):
        v, se, _, _ = orc.libor(1 << 17, L0=0.3, lam=lam, seed=3)
        assert abs(v - want) < 3 * se, (lam, v, want, se)


def pin_heston_running_val2():
    """MCAsianOptions.cpp:184-199 with xi = 0, X = 0 and all normals zero:
    every path is the same deterministic average A, each 64-path chunk adds A
    to the running val, and val2 += val^2 adds (kA)^2 after chunk k.  After c
    chunks (N = 64c paths) val = cA and val2 = A^2 c(c+1)(2c+1)/6, so the
    sample's confidence is e^{-rT} 1.96 A sqrt((N c(c+1)(2c+1)/6 - c^2) /
    (N(N-1))) / sqrt(N), while the price is e^{-rT} A and a per-path standard
    error would be 0.  With v0 = theta the variance stays at theta and A =
    S0 mean_k e^{(r - theta/2) k dt}; with v0 != theta it follows
    v' = v + kappa dt (theta - v)."""
    n_steps, c, r = 6, 5, .03
    N = 64 * c
    S, T = np.array([100.0, 90.0]), np.array([1.0, 0.5])
    v0, kap, th = np.array([.04, .09]), np.array([2.0, 3.0]), np.array([.04, .05])
    z = np.zeros((2, c, n_steps, 64))
    res, conf = orc.heston_buffer(S, np.zeros(2), T, v0, kap, th, np.zeros(2), r, -.7, N, n_steps, z, z,
                                  quirks=orc.Q_RUNNING_VAL2)
    for o in range(2):
        dt = T[o] / n_steps
        v, s, avg = v0[o], S[o], 0.0
        for _ in range(n_steps):
            s *= math.exp((r - .5 * v) * dt)
            v += kap[o] * dt * (th[o] - v)
            avg += s
        A = avg / n_steps
        if o == 0:
            assert A == pytest.approx(S[0] * np.mean(np.exp((r - th[0] / 2) * dt * np.arange(1, n_steps + 1))),
                                      rel=1e-14)
        disc = math.exp(-r * T[o])
        want = disc * 1.96 * A * math.sqrt((N * c * (c + 1) * (2 * c + 1) / 6 - c * c) / (N * (N - 1))) / math.sqrt(N)
        assert res[o] == pytest.approx(disc * A, rel=1e-13)
        assert conf[o] == pytest.approx(want, rel=1e-12)


def pin_geometric_asian_discrete():
    """The discretely monitored geometric-average call (t_k = kT/N, k = 1..N)
    against an exact-GBM Monte Carlo of 2^20 paths, within 4 SE, for N = 4
    and 12: log G is normal with mean log S + (r - sigma^2/2) T (N+1)/(2N)
    and variance sigma^2 T (N+1)(2N+1)/(6N^2).  The continuous-average
    variance sigma^2 T/3 misses by 25 SE (N = 4), the continuous mean by 60."""
    S, K, T, r, sig = 100.0, 100.0, 2.0, .05, .4
    rng = np.random.default_rng(23)
    for N in (4, 12):
        dt = T / N
        logS = np.cumsum((r - .5 * sig * sig) * dt + sig * math.sqrt(dt) * rng.standard_normal((1 << 20, N)),
                         axis=1)
        pay = math.exp(-r * T) * np.maximum(S * np.exp(logS.mean(axis=1)) - K, 0)
        se = pay.std() / math.sqrt(pay.size)
        assert abs(orc.geometric_asian_call(S, K, T, r, sig, N) - pay.mean()) < 4 * se


def test_lsmc_ref_two_dates_closed_form():
    pin_lsmc_ref_two_dates()


def test_lsmc_ref_lag_and_ridge():
    pin_lsmc_ref_lag_and_ridge()


def test_libor_spot_measure_drift():
    pin_libor_spot_measure_drift()


def test_heston_running_val2_confidence():
    pin_heston_running_val2()


def test_geometric_asian_discrete_monitoring():
    pin_geometric_asian_discrete()


def test_sfmt19937_sample_scaled_draws_against_mkl_fixture():
    # The European sample draws vsRngGaussian(ICDF, ..., MuByT, VBySqrtT)
    # (MonteCarloInsideBlockingDP.cpp:249): MKL applies the (float) mean and
    # deviation in float arithmetic -- its scaled block equals fl(m) + fl(s) z
    # of its own unscaled normals to one ulp -- so what separates the engine's
    # FP64 inverse CDF from MKL is MKL's float ICDF alone: within 6e-5 |z|
    # (its error grows in the tails, 2e-4 at z = -4.18).
    from scipy.special import ndtri
    d = _mkl()
    blocks, mean, sd = d["sfmt777_sample_blocks"], d["sfmt777_sample_mean"], d["sfmt777_sample_sd"]
    m0, s0 = np.float32(mean[0]), np.float32(sd[0])
    own = (m0 + s0 * d["sfmt777_icdf_float"][:2048]).astype(np.float32)
    assert np.max(np.abs(blocks[0] - own)) <= 6e-8
    z = ndtri(orc.sfmt19937_icdf_uniforms(777, 3 * 2048)).reshape(3, 2048)
    exact = mean[:, None] + sd[:, None] * z
    assert np.all(np.abs(blocks - exact) <= sd[:, None] * 6e-5 * np.maximum(1.0, np.abs(z)) + 1e-7)
