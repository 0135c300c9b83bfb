"""Build librmblast_hip.so and the rmblastn front end for gfx950, in-tree.

Outputs go to rmblast_amd/_build/.  Run as `python -m rmblast_amd.build`.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RMB_OFFLOAD_ARCH", "gfx950")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]
LIB_SOURCES = ["rmb_kernels.hip", "rmb_prep.hip", "rmb_pipeline.cpp", "rmb_replay.cpp", "rmb_dust.cpp", "rmb_ka.cpp", "rmb_report.cpp", "triage_kernels.hip",
               "triage.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library():
    os.makedirs(OUT, exist_ok=True)
    lib = os.path.join(OUT, "librmblast_hip.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, *COMMON, "-shared", "-o", tmp,
          *[os.path.join(CSRC, s) for s in LIB_SOURCES], "-lpthread"])
    os.replace(tmp, lib)
    return lib


def build_cli():
    exe = os.path.join(OUT, "rmblastn")
    _run([HIPCC, *COMMON, "-o", exe, os.path.join(CSRC, "rmblastn.cpp"),
          "-L" + OUT, "-lrmblast_hip", "-Wl,-rpath,$ORIGIN"])
    return exe


def build_oracle():
    """CPU oracle (test infrastructure) under oracle/_build."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-j8"])


def main():
    build_library()
    build_cli()
    build_oracle()


if __name__ == "__main__":
    sys.exit(main())
