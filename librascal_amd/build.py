"""Build librascal_mi355x.so (HIP, gfx950) in-tree.

    python -m librascal_amd.build          # or librascal_amd.build.build()

Each source is compiled by hipcc for gfx950 only, then linked into
librascal_amd/_lib/librascal_mi355x.so, which the Python package and C/C++
callers load.  Compiles every time it is called (object files are not
cached), in parallel.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
BUILD_DIR = os.path.join(HERE, "_build")
LIB_NAME = "librascal_mi355x.so"
LIB_PATH = os.path.join(LIB_DIR, LIB_NAME)

ARCH = "gfx950"
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
# geometry of the neighbour list must round like separate multiply/add
SOURCE_FLAGS = {"neighbours.hip": ["-ffp-contract=off"]}
SOURCES = ["hypers.cc", "radial_host.cc", "neighbours.hip", "soap_kernels.hip",
           "gap_kernels.hip", "gap_mfma.hip", "gap_f64.hip", "context.cc"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def _compile(src):
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [_hipcc()] + COMMON + SOURCE_FLAGS.get(src, []) + lang + ["-c", os.path.join(CSRC, src),
                                                                   "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature
    err = "\n".join(ln for ln in res.stderr.splitlines()
                    if "packed-fp32-ops" not in ln and ln.strip())
    if res.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, err))
    return obj, err


def build(verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = []
    for obj, err in results:
        objs.append(obj)
        if verbose and err:
            print(err, file=sys.stderr)
    tmp = LIB_PATH + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + res.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
