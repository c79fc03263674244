"""Jump-ahead polynomials of the sample generators (CPU, no GPU).

fsmc_jump_poly returns g = x^J mod phi for MT19937 (steps of one 32-bit word)
and SFMT19937 (steps of one 128-bit word).  The identity the GPU skip-ahead
rests on is checked here on the oracle's own sequential streams: word k of
the window J steps ahead equals the XOR of the words i + k over the set
coefficients g_i.  The tempering of MT19937 is linear, so the identity holds
for its outputs as for its raw words.
"""
import ctypes

import numpy as np
import pytest

import mcengine
import oracle as orc


@pytest.fixture(scope="module")
def lib():
    if not __import__("os").path.exists(mcengine.LIB_PATH):
        import build
        build.build()
    L = mcengine.lib()
    L.fsmc_jump_poly.restype = ctypes.c_int
    L.fsmc_jump_poly.argtypes = [ctypes.c_int, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]
    return L


def _poly(lib, gen, J):
    w = np.zeros(312, dtype=np.uint64)
    deg = lib.fsmc_jump_poly(gen, J, w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 312)
    assert deg in (19937, 19968)
    bits = np.unpackbits(w.view(np.uint8), bitorder="little")
    return deg, np.nonzero(bits)[0]


def _window_xor(seq, idx, n):
    """XOR over i in idx of seq[i + k], k = 0..n-1 (seq: uint32 array)."""
    acc = np.zeros(n, dtype=np.uint32)
    for i0 in range(0, len(idx), 512):
        block = seq[idx[i0:i0 + 512, None] + np.arange(n)[None, :]]
        acc ^= np.bitwise_xor.reduce(block, axis=0)
    return acc


@pytest.mark.parametrize("J", [1, 623, 624, 1000, 19937, 123457])
def test_mt19937_jump_identity(lib, J):
    deg, idx = _poly(lib, 0, J)
    assert deg == 19937 and idx.max() < deg
    out = orc.mt19937_bits(42, 0, int(idx.max()) + 624 + 1).astype(np.uint32)
    want = orc.mt19937_bits(42, J, 624).astype(np.uint32)
    assert np.array_equal(_window_xor(out, idx, 624), want)


@pytest.mark.parametrize("J", [1, 155, 156, 5000, 19968, 30011])
def test_sfmt19937_jump_identity(lib, J):
    deg, idx = _poly(lib, 1, J)
    assert deg == 19968 and idx.max() < deg
    n_words = int(max(idx.max() + 156, J + 156)) + 1
    out = orc.sfmt19937_bits(777, 4 * n_words).astype(np.uint32).reshape(-1, 4)
    for lane in range(4):
        assert np.array_equal(_window_xor(out[:, lane], idx, 156), out[J:J + 156, lane])


def test_bad_arguments(lib):
    w = np.zeros(312, dtype=np.uint64)
    p = w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    assert lib.fsmc_jump_poly(2, 5, p, 312) == -1
    assert lib.fsmc_jump_poly(0, 5, p, 311) == -1
    assert lib.fsmc_jump_poly(0, 5, None, 312) == -1
