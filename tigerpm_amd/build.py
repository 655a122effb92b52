"""Builds libtigerpm_hip.so in-tree with hipcc for gfx950.

Usage: python -m tigerpm_amd.build   (or tigerpm_amd.build.build()).
Every .hip source under csrc/ is compiled to an object in parallel and
linked against rocFFT.  Packed-FP32 instructions are disabled for all device
code (-packed-fp32-ops).
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtigerpm_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

CXXFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + CXXFLAGS + ["-DTPM_EXPORT", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-8000:]))
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + [
        "-L" + os.path.join(ROCM, "lib"), "-lrocfft", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    os.replace(tmp, LIB)
    build_host()
    if verbose:
        print("built", LIB, REF_LIB, DRIVER)
    return LIB


HOST = os.path.join(HERE, "host")
REF_LIB = os.path.join(HERE, "libtigerpm_ref.so")
DRIVER = os.path.join(HERE, "bin", "tigerpm")


def build_host():
    """The reference-named C++ layer (libtigerpm_ref.so) and the --test=
    driver, host-only code linked against libtigerpm_hip.so."""
    inc = ["-I" + os.path.join(ROOT, "include")]
    rpath = "-Wl,-rpath,$ORIGIN:$ORIGIN/.."
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall"] + inc + [
        os.path.join(HOST, "tigerpm_ref.cpp"), "-o", REF_LIB, "-L" + HERE, "-ltigerpm_hip", rpath]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall"] + inc + [
        os.path.join(HOST, "tigerpm_main.cpp"), "-o", DRIVER, "-L" + HERE, "-ltigerpm_ref", "-ltigerpm_hip", rpath]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
