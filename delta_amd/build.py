"""Build recipe of libdelta_hip.so (hipcc, gfx950) and the host-side check
library.  Usage: python -m delta_amd.build  (or delta_amd.build.build())."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdelta_hip.so")
HOSTCHECK = os.path.join(HERE, "libdh_hostcheck.so")

HIP_SOURCES = ["runtime.hip", "tri_pairs.hip", "mixed.hip", "detect.hip", "dynamics.hip",
               "broadphase.hip", "engine.hip", "learned.hip", "zoltan.hip",
               "ispc.hip", "adaptive.hip", "tuning.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # exact paths are binary64 without fused multiply-add (reference parity)
    "-ffp-contract=off",
    # packed FP32 VALU ops are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


def build_lib(sources=None):
    srcs = [os.path.join(CSRC, s) for s in (sources or HIP_SOURCES)]
    out_tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", out_tmp] + srcs
    _run(cmd)
    os.replace(out_tmp, LIB)


def build_hostcheck():
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", HOSTCHECK,
           os.path.join(CSRC, "host_check.cpp")]
    _run(cmd)


def build():
    build_lib()
    build_hostcheck()


if __name__ == "__main__":
    build()
    sys.exit(0)
