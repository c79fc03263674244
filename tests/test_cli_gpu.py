"""Sample-compatible command lines (bin/<sample>_{dp,sp}): the lines the
samples' MPTest getresults.sh scripts grep are printed in the samples'
formats, and the printed results are right (the European sample's own
Black-Scholes check; the American sample's prices equal the engine's on the
sample's stream)."""
import os
import re
import subprocess

import pytest

import mcengine as mc

pytestmark = pytest.mark.gpu


def _run(name, *args):
    exe = os.path.join(mc.HERE, "bin", name)
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout, r.stderr


@pytest.mark.parametrize("build", ["dp", "sp"])
def test_european_command_line(build):
    out, _ = _run(f"european_{build}", "4", "1k", "32k", "8k", "1")
    # MPTest: grep Elapsed | awk '{s+=$4}'
    line = [l for l in out.splitlines() if "Elapsed" in l][0]
    assert float(line.split()[3]) > 0
    assert re.search(r"Opt/sec\s+= [0-9.]+", out)
    assert "Test passed" in out
    l1 = float(re.search(r"L1_Norm\s+= ([0-9.eE+-]+)", out).group(1))
    assert l1 < 5e-3  # 32768 paths per option against Black-Scholes (measured 1.4e-3)


def test_european_argument_checks():
    exe = os.path.join(mc.HERE, "bin", "european_sp")
    r = subprocess.run([exe, "3", "1k", "32k", "8k"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "multiple of 4" in r.stderr
    r = subprocess.run([exe, "4", "1k", "32k", "5000"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0


@pytest.mark.parametrize("build", ["dp", "sp"])
def test_asian_command_line(build):
    out, err = _run(f"asian_{build}")
    # the sample's "%8.4f" line; MPTest's grep adds three spaces after "-",
    # which only matches rates below 10 options/s (one CPU core), so the GPU
    # line is found without them; awk's $8 is still the rate
    line = [l for l in out.splitlines() if "Computation rate MC by time Step -" in l][0]
    rate = float(line.split()[7].rstrip("."))
    assert rate > 0 and float(err.split()[-1]) == pytest.approx(rate, rel=1e-3)


@pytest.mark.parametrize("build", ["dp", "sp"])
def test_american_command_line(engine, build):
    out, _ = _run(f"american_{build}")
    # MPTest: grep Computation | awk '{s+=$5}'
    line = [l for l in out.splitlines() if "Computation" in l][0]
    assert float(line.split()[4]) > 0
    eu = float(re.search(r"European Put Option Price: ([0-9.eE+-]+)", out).group(1))
    am = float(re.search(r"American Put Option Price: ([0-9.eE+-]+)", out).group(1))
    # the printed prices are those of the last loop (50) on the sample's own
    # stream (its overlapping MRG32k3a streams correlate paths, so they are not
    # iid estimates of the Black-Scholes put); 9 significant digits printed
    # of a REAL (float in the sp build: rounding to 6e-8)
    prec = mc.FP64 if build == "dp" else mc.FP32
    m = mc.mode(prec, mc.RNG_MRG32K3A, seed=12345, stream=50)
    ram, reu = engine.american_lsmc_ref(524288, 256, 100.0, 110.0, .25, 2.0, m)
    tol = 2e-8 if build == "dp" else 1e-7
    assert am == pytest.approx(ram, rel=tol) and eu == pytest.approx(reu, rel=tol)
