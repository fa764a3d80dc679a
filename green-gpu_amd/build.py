"""Build libgreen_gpu_amd.so (gfx950) in-tree with hipcc.

    python green-gpu_amd/build.py            # build
    python green-gpu_amd/build.py --clean    # remove objects and library

Objects go to green-gpu_amd/_build/, the library to green-gpu_amd/.
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
LIB = os.path.join(HERE, "libgreen_gpu_amd.so")
SITE_BENCH = os.path.join(HERE, "gemm_site_bench")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["zgemm.hip", "ozaki.hip", "kernels.hip", "engine.hip", "capi.cpp"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-result"]
# Device code: no packed FP32 instructions (hipcc would otherwise form
# v_pk_*_f32 from scalar float code).
DEVICE = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + (DEVICE if src.endswith(".hip") else []) + \
        ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{msgs}")
    return obj, msgs


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(5, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    for _, msgs in results:
        if verbose and msgs.strip():
            print(msgs)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    # the GEMM call-site benchmark, linked against the library just built
    cmd = [HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}", os.path.join(CSRC, "gemm_site_bench.cpp"),
           "-o", SITE_BENCH, "-L" + HERE, "-lgreen_gpu_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("gemm_site_bench build failed:\n" + r.stdout + r.stderr)
    return LIB


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)
    for f in (LIB, SITE_BENCH):
        if os.path.exists(f):
            os.remove(f)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--clean", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    if a.clean:
        clean()
    else:
        print(build(verbose=a.verbose))
    sys.exit(0)
