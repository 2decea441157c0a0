"""Build recipe of libcmi.so (the C-ABI library with the gfx950 HIP kernels).

    python -m parallelcmacionize_amd.build

Compiles every parallelcmacionize_amd/csrc/*.hip for gfx950 with hipcc and
links them into parallelcmacionize_amd/libcmi.so (in-tree).
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libcmi.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CMI_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the FP64 path must follow the reference arithmetic
# without fused multiply-adds; FP32 code asks for FMAs explicitly.
# -target-feature -packed-fp32-ops: no packed FP32 instructions.
CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-ffp-contract=off", "-fno-fast-math", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            if err.strip():
                sys.stderr.write(err)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    build_host_programs()
    return LIB


ROOT = os.path.dirname(PKG)
HOST_PROGRAMS = {
    # C++ programs that use the host interface (host/cmi.hpp) over libcmi.so
    os.path.join(ROOT, "tests", "cpp", "test_facade"):
        os.path.join(ROOT, "tests", "cpp", "test_facade.cpp"),
}


def build_host_programs():
    for out, src in HOST_PROGRAMS.items():
        rel = os.path.relpath(PKG, os.path.dirname(out))
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-o", out, src,
               "-L" + PKG, "-lcmi", "-Wl,-rpath,$ORIGIN/" + rel]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("g++ failed:\n%s\n%s" % (" ".join(cmd), r.stderr))


if __name__ == "__main__":
    print(build(verbose=True))
