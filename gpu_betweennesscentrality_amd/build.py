"""Build the bcx native library, the three drop-in executables and the oracle.

Everything is built in-tree so that the shared objects travel with the
repository snapshot:
  gpu_betweennesscentrality_amd/libbcx.so      HIP kernels (gfx950) + host C++ + C ABI
  gpu_betweennesscentrality_amd/bin/{QBC,MHLQBC,RQ}
  oracle/liboracle.so                          CPU oracle (test infrastructure)
  oracle/_ref/tc_ref                           reference host index dumper, only
                                               where the reference checkout exists
Usage: python -m gpu_betweennesscentrality_amd.build [--no-oracle]
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BCX_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # Packed FP32 instructions stay disabled for every HIP source.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-Wall", "-Wno-unused-function",
             f"-I{INCLUDE}", f"-I{CSRC}"]
HOST_SOURCES = ["graph_io.cpp", "index_build.cpp", "bcx_api.cpp"]
HIP_SOURCES = ["bcx_device.hip", "bcx_sweep.hip"]
CLIS = {"MHLQBC": "cli_mhlqbc.cpp", "QBC": "cli_qbc.cpp", "RQ": "cli_rq.cpp",
        "GraphConstract": "cli_graphconstract.cpp", "TC_index": "cli_tc_index.cpp",
        "TC_query": "cli_tc_query.cpp"}


def _run(cmd: list[str]) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def _stale(out: Path, deps: list[Path]) -> bool:
    if not out.exists():
        return True
    t = out.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def build_library(force: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    headers = list(CSRC.glob("*.hpp")) + list(INCLUDE.glob("*.h"))
    objs = []
    for src in HOST_SOURCES:
        obj = BUILD / (src + ".o")
        if force or _stale(obj, [CSRC / src] + headers):
            _run(["g++", *CXX_FLAGS, "-c", CSRC / src, "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        if force or _stale(obj, [CSRC / src] + headers):
            _run([HIPCC, *HIP_FLAGS, "-c", CSRC / src, "-o", obj])
        objs.append(obj)
    lib = PKG / "libbcx.so"
    tmp = PKG / "libbcx.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lgomp"])
    os.replace(tmp, lib)
    bindir = PKG / "bin"
    bindir.mkdir(exist_ok=True)
    for exe, src in CLIS.items():
        _run(["g++", *CXX_FLAGS, CSRC / src, "-o", bindir / exe, f"-L{PKG}", "-lbcx",
              "-Wl,-rpath,$ORIGIN/.."])
    return lib


def build_oracle() -> Path:
    odir = ROOT / "oracle"
    lib = odir / "liboracle.so"
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-Wall", odir / "oracle.c", "-o", lib,
          "-lm"])
    ref = os.environ.get("BCX_REFERENCE_DIR", "/root/reference")
    if Path(ref, "src", "TreeCount", "TreeDecomp.cpp").exists():
        _run(["sh", odir / "build_ref.sh"])
    return lib


def main(argv: list[str]) -> int:
    build_library(force="--incremental" not in argv)
    if "--no-oracle" not in argv:
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
