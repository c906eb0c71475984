"""Builds libfame6d.so (HIP, gfx950) next to this file.

Each csrc/*.hip file is compiled by hipcc to an object in parallel, then the
objects are linked into one shared library with the C ABI of include/fame6d.h.
Usage: python build_lib.py [--jobs J] [--debug]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfame6d.so")
ARCH = "gfx950"


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags(debug=False):
    f = [
        "-std=c++17",
        "-O1" if debug else "-O3",
        f"--offload-arch={ARCH}",
        "-fPIC",
        # packed FP32 instructions are kept out of every kernel
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I" + os.path.join(ROOT, "include"),
        "-Wno-unused-command-line-argument",
    ]
    return f


def compile_one(src, obj, debug):
    cmd = [hipcc()] + flags(debug) + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (os.path.basename(src), p.stderr[-8000:]))
    return obj


def build(jobs=None, debug=False, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(BUILD, s[:-4] + ".o") for s in srcs]
    jobs = jobs or min(len(srcs), max(1, os.cpu_count() or 1), 8)
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(compile_one, os.path.join(CSRC, s), o, debug) for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stderr[-8000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--debug", action="store_true")
    a = ap.parse_args()
    try:
        build(a.jobs, a.debug)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
