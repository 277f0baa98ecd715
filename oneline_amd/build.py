"""Builds the native parts of oneline_amd in-tree.

* ``liboneline_hip.so``  -- the HIP engine + C ABI (hipcc, gfx950)
* ``libol_datagen.so``   -- the synthetic workload generator (gcc)

Run ``python -m oneline_amd.build`` (or ``__graft_entry__.build()``).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")

HIP_SOURCES = ["api.cpp", "scan.hip", "tree.hip", "search.hip", "calib.hip", "serve.cpp"]
HIP_LIB = os.path.join(HERE, "liboneline_hip.so")
GEN_LIB = os.path.join(HERE, "libol_datagen.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_datagen():
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
          "-o", GEN_LIB, os.path.join(CSRC, "datagen.c"), os.path.join(CSRC, "profiles.c"), "-lm", "-lpthread"])


def build_hip():
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = HIP_LIB + ".tmp"
    _run([hipcc] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs + ["-lpthread"])
    os.replace(tmp, HIP_LIB)


def build_all():
    build_datagen()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "datagen": build_datagen}[w]()
