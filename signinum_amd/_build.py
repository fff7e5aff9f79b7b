"""Builds libsigninum_j2k_hip.so in-tree for gfx950.

Host C++ (codestream parsing, plan building, the test-input encoder) is
compiled with g++; the HIP runtime and kernels with hipcc.  Everything is
linked into one shared library next to this file.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsigninum_j2k_hip.so")
ARCH = os.environ.get("SGN_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["container.cpp", "geometry.cpp", "parse.cpp", "plan.cpp", "encode.cpp", "encode_ht_refine.cpp", "encode_classic.cpp", "encode_image.cpp",
                "api_encode.cpp", "api_host.cpp"]
HIP_SOURCES = ["runtime.hip"]


def _hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build")


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + res.stdout + res.stderr)
    return res


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    common = ["-O3", "-std=c++17", "-fPIC", "-Wall"]
    jobs = []
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        jobs.append(["g++"] + common + ["-c", os.path.join(CSRC, "host", src), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        jobs.append([hipcc, "-x", "hip", f"--offload-arch={ARCH}"] + common +
                    ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                     "-Wno-unused-command-line-argument",
                     "-c", os.path.join(CSRC, "hip", src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for res in ex.map(_run, jobs):
            if verbose and (res.stdout or res.stderr):
                sys.stderr.write(res.stdout + res.stderr)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
