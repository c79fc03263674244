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
RNG_PHILOX, RNG_BUFFER = 0, 1
QUIRK_ASIAN_VAR_AS_VOL = 1 << 0
QUIRK_ASIAN_RUNNING_VAL2 = 1 << 1
QUIRK_LIBOR_INPLACE_ADJ = 1 << 2

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_VP = ctypes.c_void_p


class FsmcError(RuntimeError):
    def __init__(self, status, where):
        detail = _lib.fsmc_last_error().decode() if _lib is not None else ""
        name = _lib.fsmc_status_string(status).decode() if _lib is not None else str(status)
        super().__init__(f"{where}: {name} ({status}) {detail}")
        self.status = status


class Mode(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("rng", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("stream", ctypes.c_uint32), ("quirks", ctypes.c_uint32)]


class GbmParams(ctypes.Structure):
    _fields_ = [("n_options", ctypes.c_int), ("n_paths", ctypes.c_int), ("riskfree", ctypes.c_double),
                ("volatility", ctypes.c_double)]


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
    def gbm_european(self, S, X, T, r, sigma, n_paths, m, samples=None, greeks=False):
        n = len(S)
        p = GbmParams(n, n_paths, r, sigma)
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


# ---- asynchronous entry points on torch device tensors ------------------------
def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def gbm_european_workspace(n_options, n_paths, r, sigma, m):
    p = GbmParams(n_options, n_paths, r, sigma)
    return lib().fsmc_gbm_european_workspace(ctypes.byref(p), ctypes.byref(m))


def gbm_european_async(S, X, T, r, sigma, n_paths, m, out_res, out_conf, workspace, samples=None,
                       out_greeks=None):
    p = GbmParams(S.numel(), n_paths, r, sigma)
    _check(lib().fsmc_gbm_european_async(ctypes.byref(p), ctypes.byref(m), _ptr(S), _ptr(X), _ptr(T),
                                         _ptr(samples), _ptr(out_res), _ptr(out_conf), _ptr(out_greeks),
                                         _ptr(workspace), workspace.numel() * workspace.element_size(),
                                         _stream()), "fsmc_gbm_european_async")


def heston_asian_workspace(n_options, n_paths, n_steps, r, rho, m):
    p = HestonParams(n_options, n_paths, n_steps, r, rho)
    return lib().fsmc_heston_asian_workspace(ctypes.byref(p), ctypes.byref(m))


def heston_asian_async(params, n_paths, n_steps, r, rho, m, out_res, out_conf, workspace, z1=None, z2=None,
                       out_greeks=None):
    """params: sequence of 7 device tensors (S, X, T, v0, kappa, theta, xi)."""
    p = HestonParams(params[0].numel(), n_paths, n_steps, r, rho)
    _check(lib().fsmc_heston_asian_async(ctypes.byref(p), ctypes.byref(m), *[_ptr(t) for t in params],
                                         _ptr(z1), _ptr(z2), _ptr(out_res), _ptr(out_conf), _ptr(out_greeks),
                                         _ptr(workspace), workspace.numel() * workspace.element_size(),
                                         _stream()), "fsmc_heston_asian_async")
