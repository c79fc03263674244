"""Writes mkl_streams.npz: reference outputs of the MKL VSL generators the
samples call (Intel oneAPI MKL 2021.4, the runtime shipped in this image at
/opt/conda/lib/libmkl_rt.so; no header is needed, the calls go through
ctypes).  Run by hand where that library exists; the tests read only the
.npz.  Streams, as the samples create them:

  MRG32k3a seed 12345 (LiborMarketModel/mklRng.cpp:44-139,
  MonteCarloAmericanOptions/main.cpp:103): uniforms and ICDF normals from
  position 0, ICDF normals after vslSkipAheadStream(1000003);
  MT19937 seeds 42 and 43 (AsianOptionsPricing/mkl_rng.h:19,60): raw 32-bit
  outputs and BoxMuller2 normals;
  SFMT19937 seed 777 (MonteCarloInsideBlockingDP.cpp:229,249): raw 32-bit
  outputs and single-precision ICDF normals, and three consecutive blocks of
  2048 drawn as the sample draws them, vsRngGaussian(ICDF, ..., MuByT,
  VBySqrtT) with its per-option log2-domain mean and deviation (sample
  constants r = 0.06, sigma = 0.10, :71-72,160-163,238-240, at
  T = 1, 3.7, 5).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MRG32K3A, MT19937, SFMT19937 = 3 << 20, 8 << 20, 13 << 20  # VSL_BRNG_* values
UNIFORM_STD, BITS_STD, BOXMULLER2, ICDF = 0, 0, 1, 2       # VSL_RNG_METHOD_* values
N = 2048


def main():
    L = ctypes.CDLL("/opt/conda/lib/libmkl_rt.so")
    ver = ctypes.create_string_buffer(256)
    L.MKL_Get_Version_String(ver, 256)

    def stream(brng, seed, skip=0):
        s = ctypes.c_void_p()
        assert L.vslNewStream(ctypes.byref(s), brng, ctypes.c_uint(seed)) == 0
        if skip:
            assert L.vslSkipAheadStream(s, ctypes.c_longlong(skip)) == 0
        return s

    def draw(brng, seed, kind, n=N, skip=0):
        s = stream(brng, seed, skip)
        if kind == "bits":
            out = np.zeros(n, dtype=np.uint32)
            st = L.viRngUniformBits(BITS_STD, s, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        elif kind == "uniform":
            out = np.zeros(n)
            st = L.vdRngUniform(UNIFORM_STD, s, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                ctypes.c_double(0.0), ctypes.c_double(1.0))
        elif kind in ("icdf", "boxmuller2"):
            out = np.zeros(n)
            st = L.vdRngGaussian(ICDF if kind == "icdf" else BOXMULLER2, s, n,
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.c_double(0.0),
                                 ctypes.c_double(1.0))
        elif kind == "icdf_float":
            out = np.zeros(n, dtype=np.float32)
            st = L.vsRngGaussian(ICDF, s, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                 ctypes.c_float(0.0), ctypes.c_float(1.0))
        assert st == 0
        L.vslDeleteStream(ctypes.byref(s))
        return out

    def sample_blocks(seed, mean, sd, n=N):
        """Consecutive blocks of one stream, block k with (mean[k], sd[k])
        passed as the sample passes them (double arguments of a float call)."""
        s = stream(SFMT19937, seed)
        out = np.zeros((len(mean), n), dtype=np.float32)
        for k in range(len(mean)):
            assert L.vsRngGaussian(ICDF, s, n, out[k].ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                   ctypes.c_float(mean[k]), ctypes.c_float(sd[k])) == 0
        L.vslDeleteStream(ctypes.byref(s))
        return out

    log2e, r, sig = 1.4426950408889634074, 0.06, 0.10
    T = np.array([1.0, 3.7, 5.0])
    mu_by_t = log2e * (r - 0.5 * sig * sig) * T
    v_by_sqrt_t = log2e * sig * np.sqrt(T)

    np.savez_compressed(
        os.path.join(HERE, "mkl_streams.npz"),
        mkl_version=np.frombuffer(ver.value, dtype=np.uint8),
        mrg12345_uniform=draw(MRG32K3A, 12345, "uniform"),
        mrg12345_icdf=draw(MRG32K3A, 12345, "icdf"),
        mrg12345_skip1000003_icdf=draw(MRG32K3A, 12345, "icdf", n=256, skip=1000003),
        mt42_bits=draw(MT19937, 42, "bits"),
        mt42_boxmuller2=draw(MT19937, 42, "boxmuller2"),
        mt43_bits=draw(MT19937, 43, "bits"),
        mt43_boxmuller2=draw(MT19937, 43, "boxmuller2"),
        sfmt777_bits=draw(SFMT19937, 777, "bits"),
        sfmt777_icdf_float=draw(SFMT19937, 777, "icdf_float"),
        sfmt777_sample_mean=mu_by_t,
        sfmt777_sample_sd=v_by_sqrt_t,
        sfmt777_sample_blocks=sample_blocks(777, mu_by_t, v_by_sqrt_t),
    )


if __name__ == "__main__":
    main()
