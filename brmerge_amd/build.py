"""Builds brmerge_amd/libbrmerge.so (HIP for gfx950 + host C++) in-tree.

    python -m brmerge_amd.build          # or brmerge_amd.build.build()

Device code is compiled with -ffp-contract=off (products are rounded before
any add, as in the reference) and without packed FP32 instructions.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbrmerge.so")
DRIVER = os.path.join(HERE, "brm_driver")
XCHECK = os.path.join(HERE, "libbrm_xcheck.so")
COMPAT_CALLER = os.path.join(os.path.dirname(HERE), "tests", "compat", "compat_caller")

HIP_SOURCES = ["alloc.hip", "scan.hip", "flop.hip", "short.hip", "lds.hip", "long.hip", "assemble.hip", "api.hip",
               "batch.hip", "serve.hip", "bmerge.hip", "rmed.hip",
               "transpose.hip", "hash.hip", "noalloc.hip", "small.hip"]
CXX_SOURCES = ["gen.cpp"]

ARCH = os.environ.get("BRM_OFFLOAD_ARCH", "gfx950")


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-pthread", "-ffp-contract=off", "-Wall"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + out)
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        objs.append(o)
        jobs.append([hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        objs.append(o)
        jobs.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    nproc = min(len(jobs), max(1, min(16, os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(nproc) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"])
    os.replace(tmp, LIB)
    # command-line driver over the C++ compatibility header
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(os.path.dirname(HERE), "include"),
          os.path.join(CSRC, "driver.cpp"), "-o", DRIVER + ".tmp", "-L", HERE, "-lbrmerge",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + "/opt/rocm/lib"])
    os.replace(DRIVER + ".tmp", DRIVER)
    # a test caller of the reference's C++ interface, built the same way
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(os.path.dirname(HERE), "include"),
          COMPAT_CALLER + ".cpp", "-o", COMPAT_CALLER + ".tmp", "-L", HERE, "-lbrmerge",
          "-Wl,-rpath,$ORIGIN/../../brmerge_amd", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(COMPAT_CALLER + ".tmp", COMPAT_CALLER)
    # vendor cross-check (rocSPARSE), a separate test/bench library
    _run([hipcc, "-O2", "-fPIC", "-shared", "-std=c++17", os.path.join(HERE, "xcheck", "rocsparse_xcheck.cpp"),
          "-o", XCHECK + ".tmp", "-L/opt/rocm/lib", "-lrocsparse", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(XCHECK + ".tmp", XCHECK)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
