"""Builds liblym.so (gfx950) in-tree with hipcc.

    python -m tfleonardym_amd.build            # compile every HIP source, link
    python -m tfleonardym_amd.build --check    # also the CPU check library

Each translation unit compiles in its own hipcc process (in parallel); the
objects are linked into ``tfleonardym_amd/liblym.so``.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblym.so")
CHECK_LIB = os.path.join(HERE, "liblym_cpucheck.so")

SOURCES = ["api.hip", "wilson.hip", "blas.hip", "solvers.hip", "force.hip", "measure.hip", "gauge_io.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off (hipcc's host pass warns it does not
    # know the feature; that is expected)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
    "-I", os.path.join(HERE, "..", "include"),
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return p.returncode, p.stdout


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    rc, out = _run(cmd)
    return src, obj, rc, out, cmd


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    objs = []
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, out, cmd in ex.map(_compile, SOURCES):
            if rc != 0:
                failed.append((src, out, cmd))
            else:
                if verbose and out.strip():
                    print(f"[{src}]\n{out}")
                objs.append(obj)
    if failed:
        for src, out, cmd in failed:
            sys.stderr.write(f"=== {src} failed ===\n{' '.join(cmd)}\n{out}\n")
        raise RuntimeError(f"hipcc failed for {[f[0] for f in failed]}")
    tmp = LIB + ".tmp"
    rc, out = _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    if rc != 0:
        raise RuntimeError(f"link failed:\n{out}")
    os.replace(tmp, LIB)
    return LIB


def build_cpu_check():
    """Host-only build of the per-site Wilson arithmetic (same headers as the
    kernels) for CPU tests: indexing, gamma tables, transport, boundary factor."""
    src = os.path.join(CSRC, "cpucheck.cpp")
    tmp = CHECK_LIB + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++20", "-fPIC", "-shared", "-x", "hip", "--offload-host-only",
           f"--offload-arch={ARCH}", "-I", os.path.join(HERE, "..", "include"), src, "-o", tmp]
    rc, out = _run(cmd)
    if rc != 0:
        raise RuntimeError(f"cpu check build failed:\n{' '.join(cmd)}\n{out}")
    os.replace(tmp, CHECK_LIB)
    return CHECK_LIB


ASAN_BIN = os.path.join(BUILD, "asan_driver")


def build_asan_check():
    """Host-only AddressSanitizer build of the per-site Wilson and force
    arithmetic (cpucheck.cpp) with a driver that runs every variant on
    exactly sized buffers (asan_driver.cpp).  Test-only."""
    os.makedirs(BUILD, exist_ok=True)
    tmp = ASAN_BIN + ".tmp"
    cmd = [HIPCC, "-O1", "-g", "-std=c++20", "-x", "hip", "--offload-host-only", f"--offload-arch={ARCH}",
           "-Xarch_host", "-fsanitize=address", "-fno-omit-frame-pointer",
           "-I", os.path.join(HERE, "..", "include"), os.path.join(CSRC, "cpucheck.cpp"),
           "-x", "c++", os.path.join(CSRC, "asan_driver.cpp"), "-o", tmp]
    rc, out = _run(cmd)
    if rc != 0:
        raise RuntimeError(f"asan build failed:\n{' '.join(cmd)}\n{out}")
    os.replace(tmp, ASAN_BIN)
    return ASAN_BIN


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--check", action="store_true")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(verbose=a.v))
    if a.check:
        print(build_cpu_check())
