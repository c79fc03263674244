"""ctypes binding of the CPU oracle (oracle/fsmc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/ and ``__graft_entry__.smoke()``
(the results the GPU kernels are compared with) and by bench.py's
``cpu_baseline`` leg.  The product path
never imports this module.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_U32 = ctypes.POINTER(ctypes.c_uint32)


def build():
    """Compile the oracle with its Makefile (gcc, full-precision libm)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


@contextlib.contextmanager
def use(path):
    """Route every call of this module to another build of the oracle (the
    mutated copies the pin tests compile to show that each pin rejects a
    planted mistake); the default build comes back on exit."""
    global _lib
    saved = lib()
    _lib = load(path)
    try:
        yield _lib
    finally:
        _lib = saved


def load(path):
    """ctypes handle of an oracle shared library with every signature set."""
    L = ctypes.CDLL(path)
    L.orc_philox4x32_10.argtypes = [_U32, _U32, _U32]
    L.orc_box_muller.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _D, _D]
    L.orc_normals4.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.c_uint32, ctypes.c_uint32, _D]
    for name in ("orc_bs_call", "orc_geometric_asian_call"):
        getattr(L, name).restype = ctypes.c_double
    L.orc_bs_call.argtypes = [ctypes.c_double] * 5
    L.orc_bs_call_greeks.argtypes = [ctypes.c_double] * 5 + [_D]
    L.orc_geometric_asian_call.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int]
    for name in ("orc_crr_european_call", "orc_crr_american_put"):
        getattr(L, name).restype = ctypes.c_double
        getattr(L, name).argtypes = [ctypes.c_double] * 5 + [ctypes.c_int]
    L.orc_gbm_buffer.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 _D, _D, _D, _F, _D, _D]
    L.orc_gbm_philox.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 _D, _D, _D, ctypes.c_uint64, ctypes.c_uint32, _D, _D, _D]
    L.orc_gbm_buffer_cv.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _D, _D, _D, _F, _D, _D]
    L.orc_gbm_philox_cv.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    _D, _D, _D, ctypes.c_uint64, ctypes.c_uint32, _D, _D]
    L.orc_heston_buffer.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] * 2 + [_D] * 7 +
                                    [_D, _D, ctypes.c_uint32, _D, _D])
    L.orc_heston_philox.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] * 2 + [_D] * 7 +
                                    [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _D, _D, _D])
    L.orc_libor_path.restype = ctypes.c_double
    L.orc_libor_path.argtypes = [_D, _D, ctypes.c_double, ctypes.c_uint32, _D]
    L.orc_libor.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _D, ctypes.c_uint64,
                            ctypes.c_uint32, ctypes.c_uint32, _D, _D]
    L.orc_crr_bermudan_put.restype = ctypes.c_double
    L.orc_crr_bermudan_put.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int, ctypes.c_int]
    L.orc_lsmc_ref.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [_D, _D, _D, _D]
    _I = ctypes.POINTER(ctypes.c_int)
    L.orc_lsmc.argtypes = [ctypes.c_int] + [_D] * 5 + [_I, _I, ctypes.c_uint64, ctypes.c_uint32, _D, _D]
    L.orc_chol4_solve.restype = ctypes.c_int
    L.orc_chol4_solve.argtypes = [_D, _D, _D]
    L.orc_mrg32k3a_uniform.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64, _D]
    L.orc_icdf.restype = ctypes.c_double
    L.orc_icdf.argtypes = [ctypes.c_double]
    L.orc_mrg32k3a_gaussian.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64, ctypes.c_double,
                                        ctypes.c_double, _D]
    L.orc_mt19937_bits.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64,
                                   ctypes.POINTER(ctypes.c_uint32)]
    L.orc_mt19937_gaussian.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64, ctypes.c_double,
                                       ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    return L


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_D)


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_F)


# ---- generator ---------------------------------------------------------------
def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (ctypes.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    out = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, out)
    return [int(x) for x in out]


def box_muller(a, b):
    z0, z1 = ctypes.c_double(), ctypes.c_double()
    lib().orc_box_muller(a, b, ctypes.byref(z0), ctypes.byref(z1))
    return z0.value, z1.value


def normals4(seed, stream, tag, c0, c1, c2):
    out = (ctypes.c_double * 4)()
    lib().orc_normals4(seed, stream, tag, c0, c1, c2, out)
    return list(out)


# ---- analytic ----------------------------------------------------------------
def bs_call(S, X, T, r, sigma):
    return lib().orc_bs_call(S, X, T, r, sigma)


def bs_call_greeks(S, X, T, r, sigma):
    out = (ctypes.c_double * 3)()
    lib().orc_bs_call_greeks(S, X, T, r, sigma, out)
    return list(out)


def crr_european_call(S, X, T, r, sigma, steps=2048):
    return lib().orc_crr_european_call(S, X, T, r, sigma, steps)


def crr_american_put(S, K, T, r, sigma, steps=4096):
    return lib().orc_crr_american_put(S, K, T, r, sigma, steps)


def crr_bermudan_put(S, K, T, r, sigma, dates, steps_per_date):
    return lib().orc_crr_bermudan_put(S, K, T, r, sigma, dates, steps_per_date)


def geometric_asian_call(S, X, T, r, sigma, N):
    return lib().orc_geometric_asian_call(S, X, T, r, sigma, N)


# ---- GBM ---------------------------------------------------------------------
def gbm_buffer(S, X, T, r, sigma, samples):
    n_opt = len(S)
    n_paths = samples.shape[-1]
    S, pS = _d(S)
    X, pX = _d(X)
    T, pT = _d(T)
    sm, psm = _f(samples.reshape(n_opt, n_paths))
    res, pres = _d(np.zeros(n_opt))
    conf, pconf = _d(np.zeros(n_opt))
    lib().orc_gbm_buffer(n_opt, n_paths, r, sigma, pS, pX, pT, psm, pres, pconf)
    return res, conf


def gbm_philox(S, X, T, r, sigma, n_paths, seed, stream=0, greeks=False):
    n_opt = len(S)
    S, pS = _d(S)
    X, pX = _d(X)
    T, pT = _d(T)
    res, pres = _d(np.zeros(n_opt))
    conf, pconf = _d(np.zeros(n_opt))
    g, pg = _d(np.zeros((n_opt, 6)))
    lib().orc_gbm_philox(n_opt, n_paths, r, sigma, pS, pX, pT, seed, stream, pres, pconf,
                         pg if greeks else None)
    return (res, conf, g) if greeks else (res, conf)


def gbm_buffer_cv(S, X, T, r, samples):
    """The buffer-mode paths priced with S_T as control variate."""
    n_opt = len(S)
    n_paths = samples.shape[-1]
    S, pS = _d(S)
    X, pX = _d(X)
    T, pT = _d(T)
    sm, psm = _f(samples.reshape(n_opt, n_paths))
    res, pres = _d(np.zeros(n_opt))
    conf, pconf = _d(np.zeros(n_opt))
    lib().orc_gbm_buffer_cv(n_opt, n_paths, r, pS, pX, pT, psm, pres, pconf)
    return res, conf


def gbm_philox_cv(S, X, T, r, sigma, n_paths, seed, stream=0):
    """The Philox-mode paths priced with S_T as control variate."""
    n_opt = len(S)
    S, pS = _d(S)
    X, pX = _d(X)
    T, pT = _d(T)
    res, pres = _d(np.zeros(n_opt))
    conf, pconf = _d(np.zeros(n_opt))
    lib().orc_gbm_philox_cv(n_opt, n_paths, r, sigma, pS, pX, pT, seed, stream, pres, pconf)
    return res, conf


# ---- Heston ------------------------------------------------------------------
Q_VAR_AS_VOL = 1
Q_RUNNING_VAL2 = 2


def _heston_arrays(S, X, T, v0, kappa, theta, xi):
    arrs = [_d(a) for a in (S, X, T, v0, kappa, theta, xi)]
    return [a for a, _ in arrs], [p for _, p in arrs]


def heston_buffer(S, X, T, v0, kappa, theta, xi, r, rho, n_paths, n_steps, z1, z2, quirks=0):
    n_opt = len(S)
    keep, ptrs = _heston_arrays(S, X, T, v0, kappa, theta, xi)
    z1, pz1 = _d(z1)
    z2, pz2 = _d(z2)
    assert z1.size == n_opt * n_paths * n_steps and z2.size == z1.size
    res, pres = _d(np.zeros(n_opt))
    conf, pconf = _d(np.zeros(n_opt))
    lib().orc_heston_buffer(n_opt, n_paths, n_steps, r, rho, *ptrs, pz1, pz2, quirks, pres, pconf)
    return res, conf


def heston_philox(S, X, T, v0, kappa, theta, xi, r, rho, n_paths, n_steps, seed, stream=0, quirks=0,
                  greeks=False):
    n_opt = len(S)
    keep, ptrs = _heston_arrays(S, X, T, v0, kappa, theta, xi)
    res, pres = _d(np.zeros(n_opt))
    conf, pconf = _d(np.zeros(n_opt))
    g, pg = _d(np.zeros((n_opt, 6)))
    lib().orc_heston_philox(n_opt, n_paths, n_steps, r, rho, *ptrs, seed, stream, quirks, pres, pconf,
                            pg if greeks else None)
    return (res, conf, g) if greeks else (res, conf)


# ---- LIBOR -------------------------------------------------------------------
Q_LIBOR_INPLACE = 4


def libor_path(z, L0=0.05, lam=0.2, quirks=0):
    """One path; L0 is a scalar (flat curve) or an array of 80 rates."""
    z, pz = _d(z)
    curve, pc = _d(np.broadcast_to(np.asarray(L0, dtype=np.float64), (80,)))
    Lb, pLb = _d(np.zeros(80))
    v = lib().orc_libor_path(pz, pc, lam, quirks, pLb)
    return v, Lb


def libor(n_paths, L0=0.05, lam=0.2, z=None, seed=0, stream=0, quirks=0):
    """Returns (value, value_se, deltas[80], deltas_se[80])."""
    val, pval = _d(np.zeros(2))
    dl, pdl = _d(np.zeros(160))
    pz = None
    if z is not None:
        z, pz = _d(np.asarray(z).reshape(n_paths, 40))
    lib().orc_libor(n_paths, L0, lam, pz, seed, stream, quirks, pval, pdl)
    return val[0], val[1], dl[:80].copy(), dl[80:].copy()


# ---- Longstaff-Schwartz ---------------------------------------------------------
def lsmc_ref(n_paths, n_steps, spot, strike, vol, maturity, z, betas=False):
    """(American, European); with betas=True also the [n_steps][3]
    coefficients regressed at each date (row 0 unused)."""
    z, pz = _d(np.asarray(z).reshape(-1))
    assert z.size == n_paths * n_steps
    am, eu = ctypes.c_double(), ctypes.c_double()
    b, pb = _d(np.full((n_steps, 3), np.nan))
    lib().orc_lsmc_ref(n_paths, n_steps, spot, strike, vol, maturity, pz, ctypes.byref(am), ctypes.byref(eu),
                       pb if betas else None)
    return (am.value, eu.value, b) if betas else (am.value, eu.value)


def lsmc(S0, K, sigma, r, T, steps, paths, seed, stream=0):
    """Returns (price[n], se[n], european[n])."""
    n = len(S0)
    arrs = [_d(a) for a in (S0, K, sigma, r, T)]
    st = np.ascontiguousarray(steps, dtype=np.int32)
    pa = np.ascontiguousarray(paths, dtype=np.int32)
    _I = ctypes.POINTER(ctypes.c_int)
    pr, ppr = _d(np.zeros((n, 2)))
    eu, peu = _d(np.zeros(n))
    lib().orc_lsmc(n, *[p for _, p in arrs], st.ctypes.data_as(_I), pa.ctypes.data_as(_I), seed, stream, ppr,
                   peu)
    return pr[:, 0].copy(), pr[:, 1].copy(), eu


def chol4_solve(G, rhs):
    G, pG = _d(np.asarray(G).reshape(16))
    rhs, pr = _d(rhs)
    beta, pb = _d(np.zeros(4))
    ok = lib().orc_chol4_solve(pG, pr, pb)
    return beta if ok else None


def mrg32k3a_uniform(seed, offset, n):
    """Uniforms offset .. offset+n-1 of the MKL-compatible MRG32k3a stream."""
    u, pu = _d(np.zeros(n))
    lib().orc_mrg32k3a_uniform(seed, offset, n, pu)
    return u


def icdf(u):
    return lib().orc_icdf(float(u))


def mrg32k3a_gaussian(seed, offset, n, mean=0.0, stddev=1.0):
    """generateRand's normals (MRG32k3a + inverse CDF), stream order."""
    z, pz = _d(np.zeros(n))
    lib().orc_mrg32k3a_gaussian(seed, offset, n, mean, stddev, pz)
    return z


def mt19937_bits(seed, offset, n):
    """Outputs offset .. offset+n-1 of MKL's MT19937 stream for seed."""
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_mt19937_bits(seed, offset, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def mt19937_gaussian(seed, offset, n, mean=0.0, stddev=1.0):
    """Box-Muller normals of the MT19937 stream (offset, n even)."""
    z = np.zeros(n)
    lib().orc_mt19937_gaussian(seed, offset, n, mean, stddev, z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return z


# ---- SFMT19937 (the European sample's stream) ----------------------------------
# vslNewStream(VSL_BRNG_SFMT19937, seed) (MonteCarloInsideBlockingDP.cpp:229)
# is Saito and Matsumoto's SIMD-oriented Mersenne Twister (SFMT19937: 156
# 128-bit words, recursion w[i] = w[i] ^ (w[i] <<8 1 byte) ^ ((w[i+122] >> 11)
# & MSK) ^ (w[i-2] >>128 1 byte) ^ (w[i-1] << 18)) seeded by init_by_array
# ({seed}); vsRngGaussian's ICDF method maps each 32-bit output b, read as a
# signed integer, to u = int32(b) 2^-32 + 1/2.  Pure Python (small n only).
_SFMT_MSK = (0xdfffffef, 0xddfecb7f, 0xbffaffff, 0xbffffff6)
_SFMT_PARITY = (0x00000001, 0x00000000, 0x00000000, 0x13c9e684)
_M32 = 0xffffffff


def _sfmt_words(seed):
    """init_by_array({seed}) of the 624-word state, then the period check."""
    n, lag = 624, 11
    mid = (n - lag) // 2
    s = [0x8b8b8b8b] * n
    mix1 = lambda x: ((x ^ (x >> 27)) * 1664525) & _M32  # noqa: E731
    mix2 = lambda x: ((x ^ (x >> 27)) * 1566083941) & _M32  # noqa: E731
    key = [seed & _M32]
    i = 0
    for k in range(n):  # first pass: the key, then the index
        r = mix1(s[i] ^ s[(i + mid) % n] ^ s[(i + n - 1) % n])
        s[(i + mid) % n] = (s[(i + mid) % n] + r) & _M32
        r = (r + (len(key) if k == 0 else 0) + (key[k - 1] if 1 <= k <= len(key) else 0) + i) & _M32
        s[(i + mid + lag) % n] = (s[(i + mid + lag) % n] + r) & _M32
        s[i] = r
        i = (i + 1) % n
    for _ in range(n):  # second pass
        r = mix2((s[i] + s[(i + mid) % n] + s[(i + n - 1) % n]) & _M32)
        s[(i + mid) % n] ^= r
        r = (r - i) & _M32
        s[(i + mid + lag) % n] ^= r
        s[i] = r
        i = (i + 1) % n
    inner = 0
    for k in range(4):
        inner ^= s[k] & _SFMT_PARITY[k]
    for sh in (16, 8, 4, 2, 1):
        inner ^= inner >> sh
    if not inner & 1:  # fix the first parity bit found
        for k in range(4):
            bit = next((1 << b for b in range(32) if _SFMT_PARITY[k] >> b & 1), 0)
            if bit:
                s[k] ^= bit
                break
    return s


def sfmt19937_bits(seed, n):
    """The first n 32-bit outputs of the SFMT19937 stream of `seed`."""
    w = [sum(v << (32 * j) for j, v in enumerate(_sfmt_words(seed)[4 * q:4 * q + 4])) for q in range(156)]
    full = (1 << 128) - 1
    msk = sum(m << (32 * j) for j, m in enumerate(_SFMT_MSK))
    out = []
    while len(out) < n:
        r1, r2 = w[154], w[155]
        for q in range(156):
            a, b = w[q], w[(q + 122) % 156]
            b_sr = 0
            for j in range(4):  # 32-bit lanes shifted right by 11 within each lane
                b_sr |= (((b >> (32 * j)) & _M32) >> 11) << (32 * j)
            d_sl = 0
            for j in range(4):  # 32-bit lanes shifted left by 18 within each lane
                d_sl |= ((((r2 >> (32 * j)) & _M32) << 18) & _M32) << (32 * j)
            w[q] = a ^ ((a << 8) & full) ^ (b_sr & msk) ^ (r1 >> 8) ^ d_sl
            r1, r2 = r2, w[q]
        for v in w:
            out.extend((v >> (32 * j)) & _M32 for j in range(4))
    return np.array(out[:n], dtype=np.uint32)


def sfmt19937_icdf_uniforms(seed, n):
    """The uniforms vsRngGaussian's ICDF method draws from that stream."""
    return sfmt19937_bits(seed, n).view(np.int32).astype(np.float64) * 2.0 ** -32 + 0.5
