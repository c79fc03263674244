"""Builds libfsmc.so: every HIP translation unit under csrc/ (kernels + C ABI,
hipcc for gfx950) and the C++ drop-in entry points under host/dropin_*.cpp
(g++), linked into one shared library in lib/; then the sample-compatible
command lines under host/cli_*.cpp into bin/, each in a double (_dp, the
samples' -DUSE_DOUBLE) and a float (_sp) build; and lib/libfsmc_probe.so,
the instruction issue-rate probe bench.py measures the GPU with (not linked
into the pricing library).

Run directly (``python build.py``) or through ``__graft_entry__.build()``.
"""
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libfsmc.so")
PROBE_SRC = os.path.join(HERE, "probe", "issue_rates.hip")
PROBE_LIB = os.path.join(LIB_DIR, "libfsmc_probe.so")
BIN_DIR = os.path.join(HERE, "bin")
GXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{os.path.join(os.path.dirname(HERE), 'include')}"]
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 (v_pk_*_f32) is switched off for all device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", f"-I{INCLUDE}", f"-I{CSRC}"]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        # a drop-in that manages device buffers itself includes the HIP runtime
        hip = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"] if "hip/hip_runtime.h" in open(src).read() else []
        cmd = [GXX] + HOST_FLAGS + (["-O3"] if "gf2_" in src else []) + hip + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature.
    noise = "packed-fp32-ops' is not a recognized feature"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    if err.strip():
        sys.stderr.write(err + "\n")
    return obj


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(HOST, "dropin_*.cpp")) +
                  glob.glob(os.path.join(HOST, "gf2_*.cpp")))


def _cli(src, suffix, defines):
    name = os.path.basename(src)[len("cli_"):-len(".cpp")]
    out = os.path.join(BIN_DIR, f"{name}_{suffix}")
    # command lines that hold device buffers themselves call the HIP runtime
    hip = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,/opt/rocm/lib"] if "hip/hip_runtime.h" in open(src).read() else []
    cmd = [GXX] + HOST_FLAGS + defines + [src, "-o", out, f"-L{LIB_DIR}", "-lfsmc", "-Wl,-rpath,$ORIGIN/../lib"] + hip
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{r.stderr}")
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    tmp = PROBE_LIB + ".tmp"
    r = subprocess.run([HIPCC] + COMMON + ["-shared", PROBE_SRC, "-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("probe build failed:\n" + r.stderr)
    os.replace(tmp, PROBE_LIB)
    os.makedirs(BIN_DIR, exist_ok=True)
    clis = []
    for src in sorted(glob.glob(os.path.join(HOST, "cli_*.cpp"))):
        clis.append(_cli(src, "dp", ["-DUSE_DOUBLE"]))
        clis.append(_cli(src, "sp", []))
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources and {len(clis)} command lines")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
