"""Build recipe for libtdfs (host C++ + HIP for gfx950), the ulb/lb CLIs,
the CPU oracle, and (when the reference sources are present) the bliss
cross-check library under oracle/_ref/.

Run as ``python -m tdfs_amd.build`` or through ``__graft_entry__.build()``.
Every call recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "tdfs_amd"
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
ORACLE = ROOT / "oracle"
REF_BLISS = Path("/root/reference/bliss")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TDFS_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}")


def build_lib() -> Path:
    BUILD.mkdir(exist_ok=True)
    inc = [f"-I{ROOT / 'include'}", f"-I{CSRC}"]
    host_o = BUILD / "tdfs_host.o"
    dev_o = BUILD / "tdfs_device.o"
    _run(["g++", *CXX_FLAGS, *inc, "-c", str(CSRC / "tdfs_host.cpp"), "-o", str(host_o)])
    _run([HIPCC, *HIP_FLAGS, *inc, "-c", str(CSRC / "tdfs_device.hip"), "-o", str(dev_o)])
    lib = PKG / "libtdfs.so"
    tmp = PKG / "libtdfs.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", str(host_o), str(dev_o), "-o", str(tmp)])
    os.replace(tmp, lib)
    # ulb / lb command-line drivers (cu_test.cu equivalents)
    bindir = PKG / "bin"
    bindir.mkdir(exist_ok=True)
    for name, labeled in (("ulb", 0), ("lb", 1)):
        _run(["g++", *CXX_FLAGS, *inc, f"-DTDFS_CLI_LABELED={labeled}", str(CSRC / "tdfs_cli.cpp"),
              "-o", str(bindir / name), f"-L{PKG}", "-ltdfs", "-Wl,-rpath,$ORIGIN/.."])
    return lib


def build_oracle() -> Path:
    """The oracle (test infrastructure), by its recipe oracle/Makefile."""
    _run(["make", "-B", "-C", str(ORACLE), "liboracle.so"])
    return ORACLE / "liboracle.so"


def build_ref() -> Path | None:
    """bliss 0.73 compiled from the reference's own sources, plus the
    oracle's driver that returns Aut(P) as bliss + closure computes it
    (recipe: oracle/Makefile, outputs only under oracle/_ref/).  Skipped when
    the reference tree is absent."""
    if not REF_BLISS.is_dir():
        return None
    _run(["make", "-B", "-j4", "-C", str(ORACLE), "_ref/libbliss_ref.so", f"REF_BLISS={REF_BLISS}"])
    return ORACLE / "_ref" / "libbliss_ref.so"


def build_all() -> None:
    build_oracle()
    build_ref()
    build_lib()


if __name__ == "__main__":
    build_all()
    print("built", PKG / "libtdfs.so")
