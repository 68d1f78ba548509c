"""Builds the native parts in-tree:

  galn_amd/lib/libgaln_hip.so   HIP search kernels + C ABI + serving loop + .bwt builder (hipcc, gfx950)
  galn_amd/lib/libgaln_host.so  host side of `galn aln` and the CPU search engine (g++), links libgaln_hip
  galn_amd/bin/galn             the `galn` executable
  oracle/liboracle.so           the CPU oracle (test infrastructure), and oracle/_ref/ when
                                the reference tree is present (see oracle/Makefile)

Run ``python -m galn_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "galn_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
INC = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GALN_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # no packed-FP32 instructions anywhere in the device code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value"]
# -mpopcnt: the CPU engine's occurrence counts are popcounts (without it each one is a libgcc call);
# every x86-64 CPU since 2008 has the instruction
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-mpopcnt", "-Wall", "-Wno-unused-result"]


def _run(cmd: list[str], quiet_feature_warning: bool = False) -> None:
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    out = (r.stdout + r.stderr).splitlines()
    if quiet_feature_warning:
        out = [ln for ln in out if "packed-fp32-ops" not in ln]
    if r.returncode != 0:
        sys.stderr.write("\n".join(out) + "\n")
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_hip() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libgaln_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INC}", "-o", tmp, os.path.join(CSRC, "galn_search.hip"),
          os.path.join(CSRC, "galn_membench.hip"), os.path.join(CSRC, "galn_index.hip"), os.path.join(CSRC, "galn_server.cpp"),
          "-lpthread"],
         quiet_feature_warning=True)
    os.replace(tmp, out)
    return out


def build_host() -> tuple[str, str]:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    lib = os.path.join(LIB, "libgaln_host.so")
    _run(["g++", *HOST_FLAGS, "-O3", "-shared", f"-I{INC}", "-o", lib + ".tmp",
          os.path.join(CSRC, "host", "galn_host.cpp"), os.path.join(CSRC, "host", "galn_cpu.cpp"), f"-L{LIB}", "-lgaln_hip",
          "-Wl,-rpath,$ORIGIN", "-lz", "-lpthread"])
    os.replace(lib + ".tmp", lib)
    exe = os.path.join(BIN, "galn")
    _run(["g++", *HOST_FLAGS, f"-I{INC}", "-o", exe + ".tmp", os.path.join(CSRC, "host", "galn_main.cpp"),
          f"-L{LIB}", "-lgaln_host", "-lgaln_hip", "-Wl,-rpath,$ORIGIN/../lib", "-lz", "-lpthread"])
    os.replace(exe + ".tmp", exe)
    return lib, exe


def build_oracle() -> None:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all() -> None:
    build_hip()
    build_host()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built:", os.path.join(LIB, "libgaln_hip.so"), os.path.join(LIB, "libgaln_host.so"),
          os.path.join(BIN, "galn"), shutil.which("make") and "oracle")
