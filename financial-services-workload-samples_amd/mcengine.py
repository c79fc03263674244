"""ctypes binding of libfsmc.so (the engine's C ABI, include/fsmc.h).

Two layers, mirroring the header:
  * ``Engine`` wraps a context and the synchronous host-pointer entry points
    (numpy in, numpy out);
  * ``*_async`` helpers take torch tensors that already live on the GPU and
    launch on torch's current stream; torch is only the allocator and the
    stream here.

There is no fallback: if lib/libfsmc.so is missing or fails to load, every
entry point raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfsmc.so")

FSMC_OK = 0
FP32, FP64 = 0, 1
RNG_PHILOX, RNG_BUFFER, RNG_MRG32K3A = 0, 1, 2
QUIRK_ASIAN_VAR_AS_VOL = 1 << 0
QUIRK_ASIAN_RUNNING_VAL2 = 1 << 1
QUIRK_LIBOR_INPLACE_ADJ = 1 << 2
QUIRK_BS_POLY_CND = 1 << 3
LSMC_AUTO, LSMC_RESIDENT, LSMC_STREAMED = 0, 1, 2  # FSMC_LSMC_SWEEP_*
LATTICE_AMERICAN, LATTICE_PUT = 1, 2

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_VP = ctypes.c_void_p


class FsmcError(RuntimeError):
    def __init__(self, status, where, detail=None):
        if detail is None:
            detail = _lib.fsmc_last_error().decode() if _lib is not None else ""
        name = _lib.fsmc_status_string(status).decode() if _lib is not None else str(status)
        super().__init__(f"{where}: {name} ({status}) {detail}")
        self.status = status


class Mode(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("rng", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("stream", ctypes.c_uint32), ("quirks", ctypes.c_uint32)]


class GbmParams(ctypes.Structure):
    _fields_ = [("n_options", ctypes.c_int), ("n_paths", ctypes.c_int), ("riskfree", ctypes.c_double),
                ("volatility", ctypes.c_double), ("control_variate", ctypes.c_int)]


class LiborParams(ctypes.Structure):
    _fields_ = [("n_paths", ctypes.c_int), ("L0", ctypes.c_double), ("lam", ctypes.c_double),
                ("checkpoint_every", ctypes.c_int)]


class LsmcRefParams(ctypes.Structure):
    _fields_ = [("n_paths", ctypes.c_int), ("n_steps", ctypes.c_int), ("spot", ctypes.c_double),
                ("strike", ctypes.c_double), ("vol", ctypes.c_double), ("maturity", ctypes.c_double)]


REQ_GBM_EUROPEAN, REQ_HESTON_ASIAN = 0, 1


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n_paths", ctypes.c_int), ("n_steps", ctypes.c_int),
                ("S0", ctypes.c_double), ("K", ctypes.c_double), ("T", ctypes.c_double), ("r", ctypes.c_double),
                ("sigma", ctypes.c_double), ("v0", ctypes.c_double), ("kappa", ctypes.c_double),
                ("theta", ctypes.c_double), ("xi", ctypes.c_double), ("rho", ctypes.c_double)]


class Result(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("price", "delta", "vega", "rho", "price_se", "delta_se", "vega_se",
                                               "rho_se", "latency_us")]


class BatchTuning(ctypes.Structure):
    _fields_ = [("chunk_paths", ctypes.c_int), ("requests_per_launch", ctypes.c_int), ("replay_graph", ctypes.c_int),
                ("workgroups_per_cu", ctypes.c_int), ("work_queue", ctypes.c_int)]


class HestonParams(ctypes.Structure):
    _fields_ = [("n_options", ctypes.c_int), ("n_paths", ctypes.c_int), ("n_steps", ctypes.c_int),
                ("riskfree", ctypes.c_double), ("rho", ctypes.c_double)]


_lib = None


def lib():
    """Loads libfsmc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} not built; run __graft_entry__.build()")
        # A process may hold only one HIP runtime.  torch ships its own
        # libamdhip64 (soname libamdhip64.so.7) and its libtorch_hip asks for it
        # by file name, so torch must load first: libfsmc.so then binds to the
        # already-loaded runtime by soname instead of pulling in a second copy.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        L.fsmc_status_string.restype = ctypes.c_char_p
        L.fsmc_status_string.argtypes = [ctypes.c_int]
        L.fsmc_last_error.restype = ctypes.c_char_p
        L.fsmc_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(_VP)]
        L.fsmc_ctx_destroy.argtypes = [_VP]
        PM = ctypes.POINTER(Mode)
        L.fsmc_gbm_european.argtypes = [_VP, ctypes.POINTER(GbmParams), PM, _D, _D, _D, _F, _D, _D, _D]
        L.fsmc_gbm_european_workspace.restype = ctypes.c_size_t
        L.fsmc_gbm_european_workspace.argtypes = [ctypes.POINTER(GbmParams), PM]
        L.fsmc_gbm_european_async.argtypes = [ctypes.POINTER(GbmParams), PM] + [_VP] * 7 + [
            _VP, ctypes.c_size_t, _VP]
        L.fsmc_heston_asian.argtypes = [_VP, ctypes.POINTER(HestonParams), PM] + [_D] * 12
        L.fsmc_heston_asian_workspace.restype = ctypes.c_size_t
        L.fsmc_heston_asian_workspace.argtypes = [ctypes.POINTER(HestonParams), PM]
        L.fsmc_heston_asian_async.argtypes = [ctypes.POINTER(HestonParams), PM] + [_VP] * 12 + [
            _VP, ctypes.c_size_t, _VP]
        L.fsmc_heston_asian_tangent_workspace.restype = ctypes.c_size_t
        L.fsmc_heston_asian_tangent_workspace.argtypes = [ctypes.POINTER(HestonParams), PM]
        L.fsmc_heston_asian_tangent_async.argtypes = [ctypes.POINTER(HestonParams), PM] + [_VP] * 10 + [
            _VP, ctypes.c_size_t, _VP]
        L.fsmc_heston_asian_noalloc_limit.restype = ctypes.c_size_t
        L.fsmc_heston_asian_noalloc_limit.argtypes = []
        L.fsmc_heston_asian_noalloc.argtypes = [ctypes.POINTER(HestonParams), PM] + [_VP] * 12 + [_VP]
        PL = ctypes.POINTER(LiborParams)
        L.fsmc_libor_swaption_greeks.argtypes = [_VP, PL, PM, _D, _D, _D]
        L.fsmc_libor_workspace.restype = ctypes.c_size_t
        L.fsmc_libor_workspace.argtypes = [PL, PM]
        L.fsmc_libor_swaption_greeks_async.argtypes = [PL, PM, _VP, _VP, _VP, _VP, ctypes.c_size_t, _VP]
        L.fsmc_libor_normals.argtypes = [_VP, ctypes.c_int, PM, ctypes.c_double, ctypes.c_double, _D]
        L.fsmc_black_scholes.argtypes = [_VP, ctypes.c_int, PM] + [_D] * 4 + [ctypes.c_double] + [_D] * 3
        L.fsmc_lattice.argtypes = [_VP, ctypes.c_int, PM] + [_D] * 4 + [ctypes.c_double, ctypes.c_int,
                                                                       ctypes.c_uint, _D]
        L.fsmc_mrg32k3a_uniform.argtypes = [_VP, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_longlong, _D]
        L.fsmc_mrg32k3a_gaussian.argtypes = [_VP, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_longlong,
                                             ctypes.c_double, ctypes.c_double, _D]
        L.fsmc_mrg32k3a_gaussian_async.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_longlong,
                                                   ctypes.c_double, ctypes.c_double, _VP, _VP]
        _U32 = ctypes.POINTER(ctypes.c_uint32)
        L.fsmc_mt19937_create.argtypes = [_VP, ctypes.c_uint32, ctypes.POINTER(_VP)]
        L.fsmc_mt19937_bits.argtypes = [_VP, ctypes.c_longlong, _U32]
        L.fsmc_mt19937_gaussian.argtypes = [_VP, ctypes.c_longlong, ctypes.c_double, ctypes.c_double, _D]
        L.fsmc_mt19937_bits_async.argtypes = [_VP, ctypes.c_longlong, _VP, _VP]
        L.fsmc_mt19937_gaussian_async.argtypes = [_VP, ctypes.c_longlong, ctypes.c_double, ctypes.c_double, _VP,
                                                  _VP]
        L.fsmc_mt19937_destroy.argtypes = [_VP]
        L.fsmc_sfmt19937_gaussian.argtypes = [_VP, ctypes.c_uint32, ctypes.c_longlong, ctypes.c_double,
                                              ctypes.c_double, _F]
        L.fsmc_sfmt19937_gaussian_async.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_longlong,
                                                    ctypes.c_longlong, _VP, _VP, _VP, _VP]
        L.fsmc_sfmt19937_workspace.restype = ctypes.c_size_t
        L.fsmc_sfmt19937_workspace.argtypes = [ctypes.c_int, ctypes.c_longlong]
        L.fsmc_sfmt19937_gaussian_skip_async.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_longlong,
                                                         ctypes.c_longlong, _VP, _VP, _VP, _VP, ctypes.c_size_t,
                                                         _VP]
        _I = ctypes.POINTER(ctypes.c_int)
        L.fsmc_lsmc_plan_create.argtypes = [_VP, ctypes.c_int, PM] + [_D] * 5 + [_I, _I, ctypes.POINTER(_VP)]
        L.fsmc_lsmc_plan_create_sweep.argtypes = [_VP, ctypes.c_int, PM] + [_D] * 5 + [_I, _I, ctypes.c_int,
                                                                                  ctypes.POINTER(_VP)]
        L.fsmc_lsmc_plan_sweep.argtypes = [_VP]
        L.fsmc_lsmc_plan_workspace.restype = ctypes.c_size_t
        L.fsmc_lsmc_plan_workspace.argtypes = [_VP]
        L.fsmc_lsmc_plan_run.argtypes = [_VP, _VP]
        L.fsmc_lsmc_plan_results.argtypes = [_VP, _VP, _D, _D]
        L.fsmc_lsmc_plan_destroy.argtypes = [_VP]
        L.fsmc_american_lsmc.argtypes = [_VP, ctypes.c_int, PM] + [_D] * 5 + [_I, _I, _D, _D]
        L.fsmc_american_lsmc_ref.argtypes = [_VP, ctypes.POINTER(LsmcRefParams), PM, _D, _D, _D]
        L.fsmc_american_paths_ref.argtypes = ([_VP] + [ctypes.c_int] * 2 + [ctypes.c_double] * 3 +
                                              [PM, _D, ctypes.POINTER(_VP)])
        L.fsmc_batch_create.argtypes = [_VP, ctypes.POINTER(Request), ctypes.c_int, PM, ctypes.POINTER(BatchTuning),
                                        ctypes.POINTER(_VP)]
        L.fsmc_batch_run.argtypes = [_VP, _VP]
        L.fsmc_batch_results.argtypes = [_VP, _VP, ctypes.POINTER(Result)]
        L.fsmc_batch_autotune.argtypes = [_VP, ctypes.POINTER(BatchTuning)]
        L.fsmc_batch_autotune_microbatch.argtypes = [_VP, ctypes.c_int, ctypes.POINTER(BatchTuning)]
        L.fsmc_batch_tuning_get.argtypes = [_VP, ctypes.POINTER(BatchTuning)]
        L.fsmc_batch_destroy.argtypes = [_VP]
        _lib = L
    return _lib


def _check(status, where):
    if status != FSMC_OK:
        raise FsmcError(status, where)


def mode(precision=FP32, rng=RNG_PHILOX, seed=0, stream=0, quirks=0):
    return Mode(precision, rng, seed, stream, quirks)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_D)


class Engine:
    """A device context and the synchronous (host-pointer) entry points."""

    def __init__(self, device=0):
        self._ctx = _VP()
        _check(lib().fsmc_ctx_create(device, ctypes.byref(self._ctx)), "fsmc_ctx_create")

    def close(self):
        if self._ctx:
            lib().fsmc_ctx_destroy(self._ctx)
            self._ctx = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- GBM European --------------------------------------------------------
    def gbm_european(self, S, X, T, r, sigma, n_paths, m, samples=None, greeks=False, control_variate=False):
        """Returns (price[n], confidence[n]) or (price, confidence, greeks[n][6]);
        control_variate prices with S_T as control variate."""
        n = len(S)
        p = GbmParams(n, n_paths, r, sigma, 1 if control_variate else 0)
        S, pS = _d(S)
        X, pX = _d(X)
        T, pT = _d(T)
        psm = None
        if samples is not None:
            samples = np.ascontiguousarray(samples, dtype=np.float32).reshape(n, n_paths)
            psm = samples.ctypes.data_as(_F)
        res, pres = _d(np.zeros(n))
        conf, pconf = _d(np.zeros(n))
        g, pg = _d(np.zeros((n, 6)))
        _check(lib().fsmc_gbm_european(self._ctx, ctypes.byref(p), ctypes.byref(m), pS, pX, pT, psm, pres,
                                       pconf, pg if greeks else None), "fsmc_gbm_european")
        return (res, conf, g) if greeks else (res, conf)

    # -- Heston Asian --------------------------------------------------------
    def heston_asian(self, S, X, T, v0, kappa, theta, xi, r, rho, n_paths, n_steps, m, z1=None, z2=None,
                     greeks=False):
        n = len(S)
        p = HestonParams(n, n_paths, n_steps, r, rho)
        arrs = [_d(a) for a in (S, X, T, v0, kappa, theta, xi)]
        ptrs = [q for _, q in arrs]
        zs = [None, None]
        keep = []
        if z1 is not None:
            for i, z in enumerate((z1, z2)):
                z, pz = _d(np.asarray(z).reshape(-1))
                keep.append(z)
                zs[i] = pz
        res, pres = _d(np.zeros(n))
        conf, pconf = _d(np.zeros(n))
        g, pg = _d(np.zeros((n, 6)))
        _check(lib().fsmc_heston_asian(self._ctx, ctypes.byref(p), ctypes.byref(m), *ptrs, zs[0], zs[1], pres,
                                       pconf, pg if greeks else None), "fsmc_heston_asian")
        return (res, conf, g) if greeks else (res, conf)


    # -- LIBOR swaption portfolio -------------------------------------------
    def libor(self, n_paths, m, L0=0.05, lam=0.2, z=None, checkpoint_every=0):
        """Returns (value, value_se, deltas[80], deltas_se[80])."""
        p = LiborParams(n_paths, L0, lam, checkpoint_every)
        pz = None
        if z is not None:
            z, pz = _d(np.asarray(z).reshape(-1))
        val, pval = _d(np.zeros(2))
        dl, pdl = _d(np.zeros(160))
        _check(lib().fsmc_libor_swaption_greeks(self._ctx, ctypes.byref(p), ctypes.byref(m), pz, pval, pdl),
               "fsmc_libor_swaption_greeks")
        return val[0], val[1], dl[:80].copy(), dl[80:].copy()

    # -- Longstaff-Schwartz --------------------------------------------------
    def american_lsmc(self, S0, K, sigma, r, T, steps, paths, m):
        """Returns (price[n], se[n], european[n])."""
        n = len(S0)
        arrs = [_d(a) for a in (S0, K, sigma, r, T)]
        st = np.ascontiguousarray(steps, dtype=np.int32)
        pa = np.ascontiguousarray(paths, dtype=np.int32)
        _I = ctypes.POINTER(ctypes.c_int)
        pr, ppr = _d(np.zeros((n, 2)))
        eu, peu = _d(np.zeros(n))
        _check(lib().fsmc_american_lsmc(self._ctx, n, ctypes.byref(m), *[q for _, q in arrs], st.ctypes.data_as(_I),
                                        pa.ctypes.data_as(_I), ppr, peu), "fsmc_american_lsmc")
        return pr[:, 0].copy(), pr[:, 1].copy(), eu

    def lsmc_plan(self, S0, K, sigma, r, T, steps, paths, m, sweep=None):
        """sweep: None (FSMC_LSMC_SWEEP_AUTO), LSMC_RESIDENT or LSMC_STREAMED."""
        return LsmcPlan(self, S0, K, sigma, r, T, steps, paths, m, sweep)

    def american_lsmc_ref(self, n_paths, n_steps, spot, strike, vol, maturity, m, z=None):
        """The sample's own LSMC algorithm; returns (american, european)."""
        p = LsmcRefParams(n_paths, n_steps, spot, strike, vol, maturity)
        pz = None
        if z is not None:
            z, pz = _d(np.asarray(z).reshape(-1))
        am, eu = ctypes.c_double(), ctypes.c_double()
        _check(lib().fsmc_american_lsmc_ref(self._ctx, ctypes.byref(p), ctypes.byref(m), pz, ctypes.byref(am),
                                            ctypes.byref(eu)), "fsmc_american_lsmc_ref")
        return am.value, eu.value

    def american_paths_ref(self, n_paths, n_steps, spot, fwd_log2, vsdt_log2, m, z=None):
        """The sample's path matrix (perform_sims): array [n_steps + 1][n_paths]
        in the mode's precision, row 0 the spot."""
        dt = np.float64 if m.precision == FP64 else np.float32
        a = np.full((n_steps + 1, n_paths), spot, dtype=dt)
        rows = (_VP * (n_steps + 1))(*[None] + [a[i].ctypes.data for i in range(1, n_steps + 1)])
        pz = None
        if z is not None:
            z, pz = _d(np.asarray(z).reshape(-1))
        _check(lib().fsmc_american_paths_ref(self._ctx, n_paths, n_steps, spot, fwd_log2, vsdt_log2, ctypes.byref(m),
                                             pz, rows), "fsmc_american_paths_ref")
        return a

    # -- closed form and lattice ---------------------------------------------
    def black_scholes(self, S, X, T, sigma, r, m, greeks=False):
        """Returns (call[n], put[n]) or (call, put, greeks[n][3])."""
        n = len(S)
        arrs = [_d(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,))) for a in (S, X, T, sigma)]
        call, pc = _d(np.zeros(n))
        put, pp = _d(np.zeros(n))
        g, pg = _d(np.zeros((n, 3)))
        _check(lib().fsmc_black_scholes(self._ctx, n, ctypes.byref(m), *[q for _, q in arrs], r, pc, pp,
                                        pg if greeks else None), "fsmc_black_scholes")
        return (call, put, g) if greeks else (call, put)

    def lattice(self, S, X, T, sigma, r, steps, m, american=False, put=False):
        n = len(S)
        arrs = [_d(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,))) for a in (S, X, T, sigma)]
        out, po = _d(np.zeros(n))
        flags = (LATTICE_AMERICAN if american else 0) | (LATTICE_PUT if put else 0)
        _check(lib().fsmc_lattice(self._ctx, n, ctypes.byref(m), *[q for _, q in arrs], r, steps, flags, po),
               "fsmc_lattice")
        return out

    def mrg32k3a_uniform(self, seed, offset, n):
        """Positions offset .. offset+n-1 of the MKL-compatible MRG32k3a stream."""
        u = np.zeros(n)
        _check(lib().fsmc_mrg32k3a_uniform(self._ctx, seed, offset, n, u.ctypes.data_as(_D)),
               "fsmc_mrg32k3a_uniform")
        return u

    def mrg32k3a_gaussian(self, seed, offset, n, mean=0.0, stddev=1.0):
        """generateRand's normals (MRG32k3a + inverse CDF), stream order."""
        z = np.zeros(n)
        _check(lib().fsmc_mrg32k3a_gaussian(self._ctx, seed, offset, n, mean, stddev, z.ctypes.data_as(_D)),
               "fsmc_mrg32k3a_gaussian")
        return z

    def sfmt19937_gaussian(self, seed, n, mean=0.0, stddev=1.0):
        """The European sample's stream: SFMT19937 of `seed` (MKL seeding),
        inverse-CDF normals as float32."""
        z = np.zeros(n, dtype=np.float32)
        _check(lib().fsmc_sfmt19937_gaussian(self._ctx, seed, n, mean, stddev, z.ctypes.data_as(_F)),
               "fsmc_sfmt19937_gaussian")
        return z

    def mt19937(self, seed):
        """A sequential MKL-compatible MT19937 stream (each call continues it)."""
        return MT19937(self, seed)

    def libor_normals(self, n_paths, m, mean=0.0, stddev=1.0):
        z, pz = _d(np.zeros(n_paths * 40))
        _check(lib().fsmc_libor_normals(self._ctx, n_paths, ctypes.byref(m), mean, stddev, pz),
               "fsmc_libor_normals")
        return z.reshape(n_paths, 40)


class MT19937:
    """vslNewStream(VSL_BRNG_MT19937, seed) on the GPU: raw 32-bit outputs and
    Box-Muller normals, each call continuing the stream."""

    def __init__(self, engine, seed):
        self._g = _VP()
        _check(lib().fsmc_mt19937_create(engine._ctx, seed, ctypes.byref(self._g)), "fsmc_mt19937_create")

    def bits(self, n):
        out = np.zeros(n, dtype=np.uint32)
        _check(lib().fsmc_mt19937_bits(self._g, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
               "fsmc_mt19937_bits")
        return out

    def gaussian(self, n, mean=0.0, stddev=1.0):
        out = np.zeros(n)
        _check(lib().fsmc_mt19937_gaussian(self._g, n, mean, stddev, out.ctypes.data_as(_D)),
               "fsmc_mt19937_gaussian")
        return out

    def gaussian_async(self, n, out, stream=None, mean=0.0, stddev=1.0):
        _need(out, "float64", n, "out", "fsmc_mt19937_gaussian_async")
        _check(lib().fsmc_mt19937_gaussian_async(self._g, n, mean, stddev, out.data_ptr(),
                                                 stream if stream is not None else _stream()),
               "fsmc_mt19937_gaussian_async")

    def close(self):
        if self._g:
            lib().fsmc_mt19937_destroy(self._g)
            self._g = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LsmcPlan:
    """A batched LSMC plan: option table on the device, workspace owned."""

    def __init__(self, engine, S0, K, sigma, r, T, steps, paths, m, sweep=None):
        self.n = len(S0)
        self._p = _VP()
        arrs = [_d(a) for a in (S0, K, sigma, r, T)]
        st = np.ascontiguousarray(steps, dtype=np.int32)
        pa = np.ascontiguousarray(paths, dtype=np.int32)
        _I = ctypes.POINTER(ctypes.c_int)
        _check(lib().fsmc_lsmc_plan_create_sweep(engine._ctx, self.n, ctypes.byref(m), *[q for _, q in arrs],
                                                 st.ctypes.data_as(_I), pa.ctypes.data_as(_I),
                                                 LSMC_AUTO if sweep is None else sweep, ctypes.byref(self._p)),
               "fsmc_lsmc_plan_create_sweep")

    @property
    def sweep(self):
        """LSMC_RESIDENT or LSMC_STREAMED: the sweep this plan runs."""
        return lib().fsmc_lsmc_plan_sweep(self._p)

    def workspace_bytes(self):
        return lib().fsmc_lsmc_plan_workspace(self._p)

    def run(self, stream=None):
        _check(lib().fsmc_lsmc_plan_run(self._p, stream), "fsmc_lsmc_plan_run")

    def results(self, stream=None):
        pr, ppr = _d(np.zeros((self.n, 2)))
        eu, peu = _d(np.zeros(self.n))
        _check(lib().fsmc_lsmc_plan_results(self._p, stream, ppr, peu), "fsmc_lsmc_plan_results")
        return pr[:, 0].copy(), pr[:, 1].copy(), eu

    def close(self):
        if self._p:
            lib().fsmc_lsmc_plan_destroy(self._p)
            self._p = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A mixed GBM/Heston request batch served in micro-batches."""

    FIELDS = ("price", "delta", "vega", "rho", "price_se", "delta_se", "vega_se", "rho_se", "latency_us")

    def __init__(self, engine, requests, m, tuning=None):
        self.n = len(requests)
        self._reqs = (Request * self.n)(*requests)
        self._b = _VP()
        t = BatchTuning(*tuning) if tuning is not None else None
        _check(lib().fsmc_batch_create(engine._ctx, self._reqs, self.n, ctypes.byref(m),
                                       ctypes.byref(t) if t is not None else None, ctypes.byref(self._b)),
               "fsmc_batch_create")

    def run(self, stream=None):
        _check(lib().fsmc_batch_run(self._b, stream), "fsmc_batch_run")

    def results(self, stream=None):
        out = (Result * self.n)()
        _check(lib().fsmc_batch_results(self._b, stream, out), "fsmc_batch_results")
        return {f: np.array([getattr(r, f) for r in out]) for f in self.FIELDS}

    def autotune(self, requests_per_launch=None):
        """Tunes the batch for throughput, or with requests_per_launch the
        fastest configuration at that fixed micro-batch size."""
        t = BatchTuning()
        if requests_per_launch is None:
            _check(lib().fsmc_batch_autotune(self._b, ctypes.byref(t)), "fsmc_batch_autotune")
        else:
            _check(lib().fsmc_batch_autotune_microbatch(self._b, int(requests_per_launch), ctypes.byref(t)),
                   "fsmc_batch_autotune_microbatch")
        return t.chunk_paths, t.requests_per_launch, t.replay_graph, t.workgroups_per_cu, t.work_queue

    def tuning(self):
        t = BatchTuning()
        _check(lib().fsmc_batch_tuning_get(self._b, ctypes.byref(t)), "fsmc_batch_tuning_get")
        return t.chunk_paths, t.requests_per_launch, t.replay_graph, t.workgroups_per_cu, t.work_queue

    def close(self):
        if self._b:
            lib().fsmc_batch_destroy(self._b)
            self._b = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gbm_request(n_paths, S0, K, T, r, sigma):
    return Request(REQ_GBM_EUROPEAN, n_paths, 1, S0, K, T, r, sigma, 0, 0, 0, 0, 0)


def heston_request(n_paths, n_steps, S0, K, T, r, v0, kappa, theta, xi, rho):
    return Request(REQ_HESTON_ASIAN, n_paths, n_steps, S0, K, T, r, 0, v0, kappa, theta, xi, rho)


def mixed_requests(n, seed=4096):
    """The mixed serving batch: 70 % GBM European (paths ~U{8192..262144},
    S0~U[5,30], K~U[1,100], T~U[0.25,10], r = .02, sigma = .3) and 30 %
    Heston Asian (steps ~U{52..365}, paths ~U{16384..131072}, S0~U[80,120],
    K = 100, T = 1, r = .03, kappa = 2, theta = .04, xi = .3, rho = -.7,
    v0 = .04), drawn from one seed."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n):
        if rng.random() < 0.7:
            reqs.append(gbm_request(int(rng.integers(8192, 262145)), rng.uniform(5, 30), rng.uniform(1, 100),
                                    rng.uniform(.25, 10), .02, .3))
        else:
            reqs.append(heston_request(int(rng.integers(16384, 131073)), int(rng.integers(52, 366)),
                                       rng.uniform(80, 120), 100.0, 1.0, .03, .04, 2.0, .04, .3, -.7))
    return reqs


# ---- asynchronous entry points on torch device tensors ------------------------
# The C entry points take raw device addresses and cannot see a tensor's
# dtype, layout or size, so every tensor is checked here before any launch:
# on the current CUDA device, of the stated dtype, contiguous, and holding at
# least the elements the kernels read or write.  A failed check raises
# FsmcError(FSMC_ERR_INVALID_ARG) and nothing is enqueued.
ERR_INVALID_ARG = -1


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _need(t, dtype, n, name, where):
    """Validate device tensor t (dtype given as a torch dtype name) for n
    elements; None passes (optional arguments are checked by the C side)."""
    if t is None:
        return
    import torch
    want = getattr(torch, dtype)

    def bad(msg):
        raise FsmcError(ERR_INVALID_ARG, where, f"{name} {msg}")
    if not isinstance(t, torch.Tensor):
        bad(f"is {type(t).__name__}, not a torch.Tensor")
    if t.dtype != want:
        bad(f"has dtype {t.dtype}, expected {want}")
    if not t.is_contiguous():
        bad("is not contiguous")
    if t.numel() < n:
        bad(f"has {t.numel()} elements, needs {n}")
    if not t.is_cuda:
        bad("is not a CUDA (HIP) device tensor")
    if t.device.index != torch.cuda.current_device():
        bad(f"is on cuda:{t.device.index}, the current device is cuda:{torch.cuda.current_device()}")


def _workspace(ws, need, where):
    _need(ws, "uint8", need, "workspace", where)
    return ws.numel() * ws.element_size()


def gbm_european_workspace(n_options, n_paths, r, sigma, m, control_variate=False):
    p = GbmParams(n_options, n_paths, r, sigma, 1 if control_variate else 0)
    return lib().fsmc_gbm_european_workspace(ctypes.byref(p), ctypes.byref(m))


def sfmt19937_gaussian_async(seed0, n_streams, n_per_stream, out, seg=1, mean=None, stddev=None):
    """Streams seed0 .. seed0 + n_streams - 1 into the float32 tensor out
    ([n_streams * n_per_stream]); mean / stddev: float64 tensors per segment."""
    where = "fsmc_sfmt19937_gaussian_async"
    total = n_streams * n_per_stream
    _need(out, "float32", total, "out", where)
    nseg = (total + seg - 1) // seg if seg > 0 else 0
    _need(mean, "float64", nseg, "mean", where)
    _need(stddev, "float64", nseg, "stddev", where)
    _check(lib().fsmc_sfmt19937_gaussian_async(seed0, n_streams, n_per_stream, seg, _ptr(mean), _ptr(stddev),
                                               _ptr(out), _stream()), where)


def sfmt19937_workspace(n_streams, n_per_stream):
    return lib().fsmc_sfmt19937_workspace(n_streams, n_per_stream)


def sfmt19937_gaussian_skip_async(seed0, n_streams, n_per_stream, out, workspace, seg=1, mean=None, stddev=None):
    """As sfmt19937_gaussian_async, each stream spread over substreams started
    by skip-ahead; workspace: uint8 tensor of sfmt19937_workspace(...) bytes."""
    where = "fsmc_sfmt19937_gaussian_skip_async"
    total = n_streams * n_per_stream
    _need(out, "float32", total, "out", where)
    nseg = (total + seg - 1) // seg if seg > 0 else 0
    _need(mean, "float64", nseg, "mean", where)
    _need(stddev, "float64", nseg, "stddev", where)
    ws_bytes = _workspace(workspace, sfmt19937_workspace(n_streams, n_per_stream), where)
    _check(lib().fsmc_sfmt19937_gaussian_skip_async(seed0, n_streams, n_per_stream, seg, _ptr(mean), _ptr(stddev),
                                                    _ptr(out), _ptr(workspace), ws_bytes, _stream()), where)


def gbm_european_async(S, X, T, r, sigma, n_paths, m, out_res, out_conf, workspace, samples=None,
                       out_greeks=None, control_variate=False):
    where = "fsmc_gbm_european_async"
    _need(S, "float64", 1, "S", where)
    n = S.numel()
    for name, t in (("X", X), ("T", T), ("out_res", out_res), ("out_conf", out_conf)):
        _need(t, "float64", n, name, where)
    _need(out_greeks, "float64", 6 * n, "out_greeks", where)
    _need(samples, "float32", n * n_paths, "samples", where)
    p = GbmParams(n, n_paths, r, sigma, 1 if control_variate else 0)
    ws_bytes = _workspace(workspace, lib().fsmc_gbm_european_workspace(ctypes.byref(p), ctypes.byref(m)), where)
    _check(lib().fsmc_gbm_european_async(ctypes.byref(p), ctypes.byref(m), _ptr(S), _ptr(X), _ptr(T),
                                         _ptr(samples), _ptr(out_res), _ptr(out_conf), _ptr(out_greeks),
                                         _ptr(workspace), ws_bytes, _stream()), where)


def heston_asian_workspace(n_options, n_paths, n_steps, r, rho, m):
    p = HestonParams(n_options, n_paths, n_steps, r, rho)
    return lib().fsmc_heston_asian_workspace(ctypes.byref(p), ctypes.byref(m))


def _heston_check(params, n_paths, n_steps, out_res, out_conf, z1, z2, out_greeks, where):
    if len(params) != 7:
        raise FsmcError(ERR_INVALID_ARG, where, "params must be the 7 tensors S, X, T, v0, kappa, theta, xi")
    _need(params[0], "float64", 1, "S", where)
    n = params[0].numel()
    for name, t in zip(("S", "X", "T", "v0", "kappa", "theta", "xi"), params):
        _need(t, "float64", n, name, where)
    _need(out_res, "float64", n, "out_res", where)
    _need(out_conf, "float64", n, "out_conf", where)
    _need(out_greeks, "float64", 6 * n, "out_greeks", where)
    _need(z1, "float64", n * n_paths * n_steps, "z1", where)
    _need(z2, "float64", n * n_paths * n_steps, "z2", where)
    return n


def heston_asian_async(params, n_paths, n_steps, r, rho, m, out_res, out_conf, workspace, z1=None, z2=None,
                       out_greeks=None):
    """params: sequence of 7 device tensors (S, X, T, v0, kappa, theta, xi)."""
    where = "fsmc_heston_asian_async"
    n = _heston_check(params, n_paths, n_steps, out_res, out_conf, z1, z2, out_greeks, where)
    p = HestonParams(n, n_paths, n_steps, r, rho)
    ws_bytes = _workspace(workspace, lib().fsmc_heston_asian_workspace(ctypes.byref(p), ctypes.byref(m)), where)
    _check(lib().fsmc_heston_asian_async(ctypes.byref(p), ctypes.byref(m), *[_ptr(t) for t in params],
                                         _ptr(z1), _ptr(z2), _ptr(out_res), _ptr(out_conf), _ptr(out_greeks),
                                         _ptr(workspace), ws_bytes, _stream()), where)


def heston_asian_tangent_workspace(n_options, n_paths, n_steps, r, rho, m):
    p = HestonParams(n_options, n_paths, n_steps, r, rho)
    return lib().fsmc_heston_asian_tangent_workspace(ctypes.byref(p), ctypes.byref(m))


def heston_asian_tangent_async(params, n_paths, n_steps, r, rho, m, out_res, out_conf, out_greeks, workspace):
    """fsmc_heston_asian_tangent_async: price and the three Greeks by forward
    (tangent) propagation; params as heston_asian_async."""
    where = "fsmc_heston_asian_tangent_async"
    if out_greeks is None:
        raise FsmcError(ERR_INVALID_ARG, where, "out_greeks is required")
    n = _heston_check(params, n_paths, n_steps, out_res, out_conf, None, None, out_greeks, where)
    p = HestonParams(n, n_paths, n_steps, r, rho)
    ws_bytes = _workspace(workspace, lib().fsmc_heston_asian_tangent_workspace(ctypes.byref(p), ctypes.byref(m)),
                          where)
    _check(lib().fsmc_heston_asian_tangent_async(ctypes.byref(p), ctypes.byref(m), *[_ptr(t) for t in params],
                                                 _ptr(out_res), _ptr(out_conf), _ptr(out_greeks), _ptr(workspace),
                                                 ws_bytes, _stream()), where)


def heston_asian_noalloc_limit():
    """Bytes of the fixed device workspace of the allocation-free entry point."""
    return lib().fsmc_heston_asian_noalloc_limit()


def heston_asian_noalloc(params, n_paths, n_steps, r, rho, m, out_res, out_conf, z1=None, z2=None,
                         out_greeks=None):
    """fsmc_heston_asian_noalloc: as heston_asian_async, with the library's
    fixed device workspace instead of a caller one (nothing is allocated)."""
    where = "fsmc_heston_asian_noalloc"
    n = _heston_check(params, n_paths, n_steps, out_res, out_conf, z1, z2, out_greeks, where)
    p = HestonParams(n, n_paths, n_steps, r, rho)
    _check(lib().fsmc_heston_asian_noalloc(ctypes.byref(p), ctypes.byref(m), *[_ptr(t) for t in params],
                                           _ptr(z1), _ptr(z2), _ptr(out_res), _ptr(out_conf), _ptr(out_greeks),
                                           _stream()), where)


def libor_workspace(n_paths, m, L0=0.05, lam=0.2, checkpoint_every=0):
    p = LiborParams(n_paths, L0, lam, checkpoint_every)
    return lib().fsmc_libor_workspace(ctypes.byref(p), ctypes.byref(m))


def libor_async(n_paths, m, out_value, out_deltas, workspace, L0=0.05, lam=0.2, z=None, checkpoint_every=0):
    """out_value: 2 doubles, out_deltas: 160 doubles (device tensors)."""
    where = "fsmc_libor_swaption_greeks_async"
    _need(out_value, "float64", 2, "out_value", where)
    _need(out_deltas, "float64", 160, "out_deltas", where)
    _need(z, "float64", n_paths * 40, "z", where)
    p = LiborParams(n_paths, L0, lam, checkpoint_every)
    ws_bytes = _workspace(workspace, lib().fsmc_libor_workspace(ctypes.byref(p), ctypes.byref(m)), where)
    _check(lib().fsmc_libor_swaption_greeks_async(ctypes.byref(p), ctypes.byref(m), _ptr(z), _ptr(out_value),
                                                  _ptr(out_deltas), _ptr(workspace), ws_bytes, _stream()), where)
