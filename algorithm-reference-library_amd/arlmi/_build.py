"""Build libarlmi.so (HIP, gfx950) in place next to this package.

Every .hip file under ../csrc is compiled with hipcc for gfx950 and linked into
arlmi/libarlmi.so; the measurement tooling under ../probe goes into arlmi/libarlmi_probe.so.
Run as ``python -m arlmi._build`` from the package root, or call ``build()``.
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
CSRC = os.path.join(ROOT, "csrc")
PROBE_SRC = os.path.join(ROOT, "probe")
INCLUDE = os.path.join(os.path.dirname(ROOT), "include")
OBJ_DIR = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG_DIR, "libarlmi.so")
PROBE_LIB = os.path.join(PKG_DIR, "libarlmi_probe.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # multiply-add fusion everywhere except inside `#pragma clang fp contract(off)` scopes (the
    # numpy-order products of the visibility kernels); plain "fast" would ignore the pragmas
    "-ffp-contract=fast-honor-pragmas", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", INCLUDE, "-I", CSRC,
]


def _sources(folder=CSRC):
    return sorted(os.path.join(folder, f) for f in os.listdir(folder) if f.endswith(".hip"))


def _compile(src, verbose):
    obj = os.path.join(OBJ_DIR, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, res.stdout, res.stderr))
    return obj


def _link(objs, lib, verbose):
    tmp = lib + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (res.stdout, res.stderr))
    os.replace(tmp, lib)


def build(verbose=False, jobs=None):
    """Compile every HIP source and link libarlmi.so (and libarlmi_probe.so); returns the
    imaging library's path."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = _sources()
    psrcs = _sources(PROBE_SRC)
    jobs = jobs or min(len(srcs) + len(psrcs), max(1, (os.cpu_count() or 2) // 2))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs + psrcs))
    _link(objs[:len(srcs)], LIB, verbose)
    _link(objs[len(srcs):], PROBE_LIB, verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
