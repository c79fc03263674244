"""Host-side checks of the C++ drop-in entry points (include/fsmc_dropin.hpp)
that need no GPU: every declared entry point is exported under the
Itanium-mangled name a sample's object files link against, invert3_arr
inverts as the sample's cofactor formula does (utilities.cpp:90-123), and the
argument checks that run before any device work reject what the samples
reject."""
import ctypes
import os
import re

import numpy as np
import pytest

import mcengine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# The symbols a sample built against the header links to: mangled names of
# the REAL = double and REAL = float overloads (VSLStreamStatePtr is void*).
MANGLED = {
    "generateRand": ["_Z12generateRandPdiddi", "_Z12generateRandPfiffi"],
    "liborSwaptionGreeks": ["_Z19liborSwaptionGreeksPdiddiS_S_i", "_Z19liborSwaptionGreeksPfiffiS_S_i"],
    "MonteCarloTimeStep": ["_Z18MonteCarloTimeStepPdS_S_S_S_S_S_S_S_", "_Z18MonteCarloTimeStepPfS_S_S_S_S_S_S_S_"],
    "price_european": ["_Z14price_europeanPKdS0_S0_ddiiiPdS1_y", "_Z14price_europeanPKfS0_S0_ffiiiPfS1_y"],
    "perform_sims": ["_Z12perform_simsPPdPPvidd", "_Z12perform_simsPPfPPviff"],
    "invert3_arr": ["_Z11invert3_arrPA3_dS0_ii", "_Z11invert3_arrPA3_fS0_ii"],
    "price_american_lsmc": ["_Z19price_american_lsmciiddddPdS_j", "_Z19price_american_lsmciiffffPfS_j"],
    "asian_config": ["_ZN11fsmc_dropin12asian_configEv"],
    "american_config": ["_ZN11fsmc_dropin15american_configEv"],
    "european_config": ["_ZN11fsmc_dropin15european_configEv"],
}


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(mcengine.LIB_PATH):
        import build
        build.build()
    return mcengine.lib()


def _declared():
    text = open(os.path.join(ROOT, "include", "fsmc_dropin.hpp")).read()
    text = re.sub(r"//[^\n]*", "", text)
    return set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)) - {"defined"}


def test_every_declared_dropin_has_a_symbol():
    assert _declared() == set(MANGLED)


@pytest.mark.parametrize("sym", [s for v in MANGLED.values() for s in v])
def test_dropin_exported(lib, sym):
    assert hasattr(lib, sym), f"{sym} not exported"


def _invert3(lib, x, real):
    ct = ctypes.c_double if real == np.float64 else ctypes.c_float
    fn = getattr(lib, "_Z11invert3_arrPA3_dS0_ii" if real == np.float64 else "_Z11invert3_arrPA3_fS0_ii")
    fn.argtypes = [ctypes.POINTER(ct), ctypes.POINTER(ct), ctypes.c_int, ctypes.c_int]
    fn.restype = None
    x = np.ascontiguousarray(x, dtype=real)
    a = np.full((3, 3), np.nan, dtype=real)
    fn(a.ctypes.data_as(ctypes.POINTER(ct)), x.ctypes.data_as(ctypes.POINTER(ct)), 3, 3)
    return a


def test_invert3_arr_double(lib, capfd):
    rng = np.random.default_rng(5)
    for _ in range(20):
        # non-symmetric, so a transposed cofactor matrix would fail
        x = rng.standard_normal((3, 3)) + 3 * np.eye(3)
        a = _invert3(lib, x, np.float64)
        assert np.allclose(a, np.linalg.inv(x), rtol=1e-12, atol=1e-13)
        assert np.allclose(a @ x, np.eye(3), atol=1e-12)
    # the ridge-corrected moment matrix of main.cpp:244-250 (Hankel in
    # sum(1), sum(S), ..., sum(S^4) over in-the-money paths)
    s = rng.uniform(60, 110, 5000)
    m = [np.sum(s ** k) + 0.01 for k in range(5)]
    h = np.array([[m[0], m[1], m[2]], [m[1], m[2], m[3]], [m[2], m[3], m[4]]])
    a = _invert3(lib, h, np.float64)
    assert np.allclose(a @ h, np.eye(3), atol=1e-6)
    assert capfd.readouterr().out == ""


def test_invert3_arr_float(lib):
    x = np.array([[4.0, 1.0, -2.0], [0.5, 3.0, 1.0], [1.0, -1.0, 5.0]])
    a = _invert3(lib, x, np.float32)
    assert a.dtype == np.float32 and np.allclose(a, np.linalg.inv(x), rtol=1e-6, atol=1e-7)


def test_invert3_arr_singular_warns(lib, capfd):
    # |det| < 1e-8 prints the sample's warning and divides regardless
    x = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [1.0, 0.0, 1.0]])
    a = _invert3(lib, x, np.float64)
    assert capfd.readouterr().out == "Trying to invert non invertible matrix\n"
    assert not np.all(np.isfinite(a))


def test_perform_sims_requires_one_spot(lib, capfd):
    # a[0] must hold one spot in every column; otherwise nothing is written
    # (checked before any device work)
    cfg = _american_config(lib)
    saved = (cfg.num_paths, cfg.num_steps, cfg.pricing)
    cfg.num_paths, cfg.num_steps = 8, 2
    try:
        a = np.full((3, 8), 100.0)
        a[0, 5] = 101.0
        a[1:] = -1.0
        rows = (ctypes.c_void_p * 3)(*[a[i].ctypes.data for i in range(3)])
        fn = lib._Z12perform_simsPPdPPvidd
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        fn.restype = None
        fn(rows, None, 1, 0.0, 0.1)
        assert np.all(a[1:] == -1.0) and cfg.pricing == saved[2]
        assert "a[0][5]" in capfd.readouterr().err
    finally:
        cfg.num_paths, cfg.num_steps, cfg.pricing = saved


class AmericanConfig(ctypes.Structure):
    _fields_ = [("num_paths", ctypes.c_int), ("num_steps", ctypes.c_int), ("seed", ctypes.c_uint),
                ("pricing", ctypes.c_uint)]


def _american_config(lib):
    fn = lib._ZN11fsmc_dropin15american_configEv
    fn.restype = ctypes.POINTER(AmericanConfig)
    return fn().contents


def test_american_config_defaults(lib):
    c = _american_config(lib)
    # MonteCarloAmericanOptions/utilities.h:23-24, main.cpp:103
    assert (c.num_paths, c.num_steps, c.seed) == (524288, 256, 12345)


@pytest.mark.parametrize("block", [8, 48, 4096, 1000])
def test_price_european_rejects_block_lengths(lib, capfd, block):
    # MonteCarloInsideBlockingDP.cpp:165-174: at least 16, a multiple of 16,
    # at most and dividing the path count; results are NaN, no device touched
    fn = lib._Z14price_europeanPKdS0_S0_ddiiiPdS1_y
    D = ctypes.POINTER(ctypes.c_double)
    fn.argtypes = [D, D, D, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, D, D,
                   ctypes.c_ulonglong]
    fn.restype = None
    S, X, T = (np.array([v]) for v in (30.0, 20.0, 2.0))
    res, conf = np.zeros(1), np.zeros(1)
    fn(*[a.ctypes.data_as(D) for a in (S, X, T)], 0.06, 0.1, 1, 2048, block, res.ctypes.data_as(D),
       conf.ctypes.data_as(D), 777)
    assert np.isnan(res[0]) and np.isnan(conf[0])
    assert "blockLen" in capfd.readouterr().err


# The samples' own prototypes (signatures as their headers declare them,
# MKL's stream pointer being void *), compiled beside fsmc_dropin.hpp where
# the sample would include it, and linked against the library: a sample built
# this way resolves every call to the drop-ins.
SAMPLE_TUS = {
    "libor": """
        void generateRand(REAL* z, int num, REAL mean, REAL stddev, int seed = 12345);
        void liborSwaptionGreeks(REAL *z, int num, REAL mean, REAL std_dev, int seed, REAL *d_v, REAL *d_Lb,
                                 int numPaths);
        int main(int argc, char**) {
          if (argc > 99) {
            REAL z[40], v, lb;
            generateRand(z, 40, 0, 1);
            liborSwaptionGreeks(z, 40, 0, 1, 0, &v, &lb, 1);
          }
          return 0;
        }""",
    "american": """
        typedef void* VSLStreamStatePtr;
        void invert3_arr(REAL a[][3], REAL x[][3], int numRows, int numCols);
        void perform_sims(REAL **a, VSLStreamStatePtr *streams, int nthreads, REAL fwdFactor, REAL volSqrtdt);
        #include "fsmc_dropin.hpp"
        int main(int argc, char**) {
          if (argc > 99) {
            REAL x[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, y[3][3], am, eu;
            REAL* rows[2] = {nullptr, nullptr};
            invert3_arr(y, x, 3, 3);
            perform_sims(rows, nullptr, 1, 0, 0);
            price_american_lsmc(1024, 2, REAL(100), REAL(110), REAL(.25), REAL(2), &am, &eu);
          }
          return 0;
        }""",
    "asian_european": """
        #include "fsmc_dropin.hpp"
        int main(int argc, char**) {
          if (argc > 99) {
            REAL r[1], c[1], p[1] = {1};
            MonteCarloTimeStep(r, c, p, p, p, p, p, p, p);
            price_european(p, p, p, REAL(.02), REAL(.3), 1, 1024, 1024, r, c);
          }
          return 0;
        }""",
}


@pytest.mark.parametrize("real", ["double", "float"])
@pytest.mark.parametrize("sample", sorted(SAMPLE_TUS))
def test_sample_prototypes_link(lib, sample, real):
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    out_dir = os.path.join(os.path.dirname(mcengine.LIB_PATH), "..", "_build", "sample_link")
    os.makedirs(out_dir, exist_ok=True)
    src = os.path.join(out_dir, f"{sample}_{real}.cpp")
    with open(src, "w") as f:
        f.write(f"typedef {real} REAL;\n" + SAMPLE_TUS[sample] + "\n")
    exe = src[:-4]
    lib_dir = os.path.dirname(mcengine.LIB_PATH)
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), src, "-o", exe,
                        f"-L{lib_dir}", "-lfsmc", f"-Wl,-rpath,{lib_dir}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
