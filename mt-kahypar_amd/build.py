"""Builds libmtkahypar_amd.so in-tree: every HIP source for gfx950 plus the
host C++ sources, compiled in parallel with hipcc and linked into one shared
library next to this file."""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmtkahypar_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources():
    hip = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    cpp = sorted(glob.glob(os.path.join(HERE, "host", "*.cpp")))
    return hip, cpp


def compile_one(src, is_hip):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + (DEVICE + ["-x", "hip"] if is_hip else ["-x", "hip", "--offload-arch=gfx950"]) + \
        ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    hip, cpp = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(compile_one, s, True) for s in hip] + [ex.submit(compile_one, s, False) for s in cpp]
        objs = [f.result() for f in futs]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_programs()
    return LIB


def build_programs():
    """Example and test programs that link the library through its public
    headers only (C++ plugin interfaces, the C library interface in C)."""
    src = os.path.join(ROOT, "tests", "cpp")
    out = os.path.join(src, "_build")
    os.makedirs(out, exist_ok=True)
    rpath = "-Wl,-rpath,$ORIGIN/../../../mt-kahypar_amd"
    jobs = [["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
             os.path.join(src, "plugin_pipeline.cpp"), "-L" + HERE, "-lmtkahypar_amd", rpath,
             "-o", os.path.join(out, "plugin_pipeline")],
            ["gcc", "-O2", "-std=c11", "-Wall", "-I" + os.path.join(ROOT, "include"),
             os.path.join(src, "capi_partition.c"), "-L" + HERE, "-lmtkahypar_amd", rpath,
             "-o", os.path.join(out, "capi_partition")]]
    for cmd in jobs:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("program build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))


if __name__ == "__main__":
    print(build(int(sys.argv[1]) if len(sys.argv) > 1 else None))
