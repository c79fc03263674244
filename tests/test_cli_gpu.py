"""Sample-compatible command lines (bin/<sample>_{dp,sp}): the lines the
samples' MPTest getresults.sh scripts grep are printed in the samples'
formats, and the printed prices are right (Black-Scholes for the European
sample's own check and for the American sample's r = 0 European leg)."""
import math
import os
import re
import subprocess

import pytest
from scipy.special import ndtr

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
def test_american_command_line(build):
    out, _ = _run(f"american_{build}")
    # MPTest: grep Computation | awk '{s+=$5}'
    line = [l for l in out.splitlines() if "Computation" in l][0]
    assert float(line.split()[4]) > 0
    eu = float(re.search(r"European Put Option Price: ([0-9.eE+-]+)", out).group(1))
    am = float(re.search(r"American Put Option Price: ([0-9.eE+-]+)", out).group(1))
    # r = 0 in the sample: the European leg is the Black-Scholes put with r = 0
    S, K, T, v = 100.0, 110.0, 2.0, 0.25
    d1 = (math.log(S / K) + 0.5 * v * v * T) / (v * math.sqrt(T))
    bs = K * ndtr(-(d1 - v * math.sqrt(T))) - S * ndtr(-d1)
    se = 17.0 / math.sqrt(524288)  # payoff std below 17 here
    assert abs(eu - bs) < 5 * se
    # with r = 0 early exercise has no value: the sample's American estimate
    # sits at the European value up to its regression noise
    assert abs(am - eu) < 0.05 * eu
