"""C-ABI checks that need no GPU: the library loads, exports every entry
point declared in include/*.h, and rejects bad arguments before touching the
device."""
import ctypes
import glob
import os
import re

import pytest

import mcengine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(fsmc_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(mcengine.LIB_PATH):
        import build
        build.build()
    return mcengine.lib()


def test_declarations_found():
    assert len(_declared()) >= 10


@pytest.mark.parametrize("name", _declared())
def test_symbol_exported(lib, name):
    assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_status_strings(lib):
    assert lib.fsmc_status_string(0) == b"ok"
    assert lib.fsmc_status_string(-1) == b"invalid argument"
    assert lib.fsmc_status_string(-4) == b"workspace too small"


def test_invalid_arguments_rejected_without_device(lib):
    m = mcengine.mode()
    p = mcengine.GbmParams(0, 1024, 0.02, 0.3)  # zero options
    st = lib.fsmc_gbm_european_async(ctypes.byref(p), ctypes.byref(m), *([None] * 7), None, 0, None)
    assert st == -1
    assert b"n_options" in lib.fsmc_last_error()
    assert lib.fsmc_gbm_european_workspace(ctypes.byref(p), ctypes.byref(m)) == 0
    p = mcengine.GbmParams(4, 1024, 0.02, 0.3)
    bad = mcengine.mode(precision=7)
    st = lib.fsmc_gbm_european_async(ctypes.byref(p), ctypes.byref(bad), *([None] * 7), None, 0, None)
    assert st == -1 and b"precision" in lib.fsmc_last_error()
    # NULL device pointers are rejected before any launch
    st = lib.fsmc_gbm_european_async(ctypes.byref(p), ctypes.byref(m), *([None] * 7), None, 0, None)
    assert st == -1
    h = mcengine.HestonParams(2, 100, 12, 0.03, -0.7)  # buffer mode needs n_paths % 64 == 0
    mb = mcengine.mode(rng=mcengine.RNG_BUFFER)
    st = lib.fsmc_heston_asian_async(ctypes.byref(h), ctypes.byref(mb), *([None] * 12), None, 0, None)
    assert st == -1 and b"64" in lib.fsmc_last_error()
    h = mcengine.HestonParams(2, 128, 12, 0.03, -1.5)
    st = lib.fsmc_heston_asian_async(ctypes.byref(h), ctypes.byref(m), *([None] * 12), None, 0, None)
    assert st == -1 and b"rho" in lib.fsmc_last_error()
    # SFMT19937 streams: stream count, a mean without a scale, a NULL output
    assert lib.fsmc_sfmt19937_gaussian_async(777, 0, 10, 1, None, None, None, None) == -1
    assert b"n_streams" in lib.fsmc_last_error()
    buf = (ctypes.c_double * 1)()
    assert lib.fsmc_sfmt19937_gaussian_async(777, 1, 10, 1, buf, None, None, None) == -1
    assert b"stddev" in lib.fsmc_last_error()
    assert lib.fsmc_sfmt19937_gaussian_async(777, 1, 10, 1, None, None, None, None) == -1
    assert b"NULL output" in lib.fsmc_last_error()
    # control_variate is 0 or 1
    cv = mcengine.GbmParams(4, 1024, 0.02, 0.3, 2)
    assert lib.fsmc_gbm_european_workspace(ctypes.byref(cv), ctypes.byref(m)) == 0


def test_workspace_sizes(lib):
    m = mcengine.mode()
    p = mcengine.GbmParams(1024, 262144, 0.02, 0.3)
    ws = lib.fsmc_gbm_european_workspace(ctypes.byref(p), ctypes.byref(m))
    assert 0 < ws < 64 << 20
    h = mcengine.HestonParams(64, 131072, 365, 0.03, -0.7)
    ws = lib.fsmc_heston_asian_workspace(ctypes.byref(h), ctypes.byref(m))
    assert ws > 0
    # LIBOR: FP32 rebuilds the rows in registers up to lambda = 0.3 and
    # stores the 2380-slot trajectory beyond it, as FP64 always does
    for prec, lam, stored in ((mcengine.FP32, 0.2, False), (mcengine.FP32, 0.3, False),
                              (mcengine.FP32, 0.31, True), (mcengine.FP64, 0.2, True)):
        lp = mcengine.LiborParams(4096, 0.05, lam)
        ws = lib.fsmc_libor_workspace(ctypes.byref(lp), ctypes.byref(mcengine.mode(prec)))
        real = 8 if prec == mcengine.FP64 else 4
        assert ws > 0 and (ws >= 2380 * 4096 * real) == stored
    # checkpoints every 8 steps: rows 8, 16, 24, 32 (240 floats per path) on
    # top of the outputs; any other period is refused
    base = lib.fsmc_libor_workspace(ctypes.byref(mcengine.LiborParams(4096, 0.05, 0.2)),
                                    ctypes.byref(mcengine.mode(mcengine.FP32)))
    ck = lib.fsmc_libor_workspace(ctypes.byref(mcengine.LiborParams(4096, 0.05, 0.2, 8)),
                                  ctypes.byref(mcengine.mode(mcengine.FP32)))
    assert 240 * 4096 * 4 <= ck - base < 2380 * 4096 * 4
    assert lib.fsmc_libor_workspace(ctypes.byref(mcengine.LiborParams(4096, 0.05, 0.2, 4)),
                                    ctypes.byref(mcengine.mode(mcengine.FP32))) == 0


def test_bench_refuses_more_than_one_gpu():
    # one GPU is measured; --gpus N > 1 exits with a message before touching a device
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stderr


def test_batch_tuning_validated_without_device(lib):
    # fsmc_batch_create rejects bad tunings before any device work; with no
    # GPU here the context cannot be created, so check the declared layout
    t = mcengine.BatchTuning(4096, 0, 1, 2, 1)
    assert ctypes.sizeof(t) == 20 and (t.chunk_paths, t.requests_per_launch, t.replay_graph,
                                        t.workgroups_per_cu, t.work_queue) == (4096, 0, 1, 2, 1)
    # the C struct has the same five ints
    hdr = open(os.path.join(ROOT, "include", "fsmc.h")).read()
    body = hdr[hdr.index("typedef struct fsmc_batch_tuning {"):hdr.index("} fsmc_batch_tuning;")]
    assert [w for w in re.findall(r"int (\w+);", body)] == [f for f, _ in mcengine.BatchTuning._fields_]


# ---- argument checks of the torch-tensor wrappers (before any launch) --------
def _cpu(n, dtype="float64"):
    import torch
    return torch.zeros(n, dtype=getattr(torch, dtype))


def test_async_wrappers_check_tensors_before_launch(lib):
    import torch
    m = mcengine.mode()
    prm = [_cpu(4) for _ in range(7)]
    cases = [
        # wrong dtype of an output
        lambda: mcengine.heston_asian_async(prm, 1024, 8, .03, -.7, m, _cpu(4, "float32"), _cpu(4), _cpu(1 << 20,
                                                                                                        "uint8")),
        # Greeks output too small ([n][6] needed)
        lambda: mcengine.heston_asian_async(prm, 1024, 8, .03, -.7, m, _cpu(4), _cpu(4), _cpu(1 << 20, "uint8"),
                                            out_greeks=_cpu(23)),
        # a non-contiguous input
        lambda: mcengine.heston_asian_async([torch.zeros(8, dtype=torch.float64)[::2]] + prm[1:], 1024, 8, .03, -.7,
                                            m, _cpu(4), _cpu(4), _cpu(1 << 20, "uint8")),
        # six parameter tensors instead of seven
        lambda: mcengine.heston_asian_noalloc(prm[:6], 1024, 8, .03, -.7, m, _cpu(4), _cpu(4)),
        # buffer normals shorter than options x paths x steps
        lambda: mcengine.heston_asian_noalloc(prm, 64, 8, .03, -.7, mcengine.mode(rng=mcengine.RNG_BUFFER),
                                              _cpu(4), _cpu(4), z1=_cpu(4 * 64 * 8 - 1), z2=_cpu(4 * 64 * 8)),
        # a host tensor that passes every other check
        lambda: mcengine.heston_asian_noalloc(prm, 1024, 8, .03, -.7, m, _cpu(4), _cpu(4)),
        lambda: mcengine.gbm_european_async(prm[0], prm[1], prm[2], .02, .3, 1024, m, _cpu(4), _cpu(3),
                                            _cpu(1 << 20, "uint8")),
        lambda: mcengine.gbm_european_async(prm[0], prm[1], prm[2], .02, .3, 1024, m, _cpu(4), _cpu(4),
                                            _cpu(1 << 20, "uint8"), out_greeks=_cpu(24, "float32")),
        lambda: mcengine.libor_async(1024, m, _cpu(2), _cpu(159), _cpu(1 << 20, "uint8")),
        lambda: mcengine.libor_async(1024, m, _cpu(2), _cpu(160), _cpu(16, "uint8")),
        lambda: mcengine.sfmt19937_gaussian_async(5, 3, 100, _cpu(299, "float32")),
        lambda: mcengine.sfmt19937_gaussian_async(5, 3, 100, _cpu(300, "float32"), 100, _cpu(3), _cpu(2)),
    ]
    for i, call in enumerate(cases):
        with pytest.raises(mcengine.FsmcError) as e:
            call()
        assert e.value.status == -1, (i, str(e.value))


def test_noalloc_limit_and_size_rejection(lib):
    # 16 MiB; the rejection of a larger pricing happens before any device call
    assert lib.fsmc_heston_asian_noalloc_limit() == 16 << 20
    h = mcengine.HestonParams(64, 131072, 365, .03, -.7)
    m = mcengine.mode(mcengine.FP32)
    fake = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    args = [fake] * 7 + [None, None, fake, fake]
    # the forward pricing of 64 x 131072 fits (1.5 MB of partials and chunk sums)
    assert lib.fsmc_heston_asian_workspace(ctypes.byref(h), ctypes.byref(m)) > (16 << 20)  # sized for Greeks
    # with Greeks its checkpoints (1.5 GB) do not: rejected with FSMC_ERR_WORKSPACE
    st = lib.fsmc_heston_asian_noalloc(ctypes.byref(h), ctypes.byref(m), *args, fake, None)
    assert st == -4 and b"16777216" in lib.fsmc_last_error()
    # invalid parameters are rejected as such
    bad = mcengine.HestonParams(0, 131072, 365, .03, -.7)
    assert lib.fsmc_heston_asian_noalloc(ctypes.byref(bad), ctypes.byref(m), *args, None, None) == -1
