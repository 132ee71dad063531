"""Build the HIP library (pubsed_amd/libsedona_gpu.so) for gfx950.

Each .hip translation unit is compiled to an object in parallel, then linked
into one shared library in-tree, so the built file travels with the source
tree.  Usage: python -m pubsed_amd.build [--jobs N]
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsedona_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SGPU_ARCH", "gfx950")

FLAGS = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
    "-munsafe-fp-atomics",
    # packed FP32 ops stay disabled (the compiler emits them from plain float
    # code otherwise); the host pass reports the feature as unknown: harmless
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    # SGPU_HIPCC_EXTRA: extra hipcc flags for a build (e.g. -gline-tables-only
    # for a profile that maps samples to source lines)
    extra = os.environ.get("SGPU_HIPCC_EXTRA", "").split()
    cmd = [HIPCC] + FLAGS + extra + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s\n%s" % (src, r.stdout, r.stderr))
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(jobs=8, verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, srcs))
    objs = []
    for obj, warn in results:
        objs.append(obj)
        if warn and verbose:
            print(warn, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_host(verbose)
    return LIB


HOST = os.path.join(HERE, "host")
DRIVER = os.path.join(HOST, "drive_grey")


def build_host(verbose=True):
    """The host C++ drop-in (transport class) and its example driver, linked
    against the HIP library with an rpath to the package directory."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-o", DRIVER + ".tmp",
           os.path.join(HOST, "drive_grey.cpp"), os.path.join(HOST, "sedona_transport.cpp"),
           os.path.join(HOST, "sedona_output.cpp"),
           "-L" + HERE, "-lsedona_gpu", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host build failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print("built", DRIVER)


if __name__ == "__main__":
    j = 8
    if "--jobs" in sys.argv:
        j = int(sys.argv[sys.argv.index("--jobs") + 1])
    build(jobs=j)
