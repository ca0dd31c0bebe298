"""Builds the native libraries of hbdia_amd in-tree.

    python hbdia_amd/build.py [--oracle] [--sanitize]
        (a script: importing the package needs the built libraries)

--sanitize builds the host library (g++ only; no device code) with
AddressSanitizer and UndefinedBehaviorSanitizer for a CPU test run:
    LD_PRELOAD=$(g++ -print-file-name=libasan.so) ASAN_OPTIONS=detect_leaks=0 \
        python -m pytest tests -m "not gpu"
A plain build afterwards restores the optimised library.

Outputs (git-ignored, shipped with the tree):
    hbdia_amd/lib/libhbdia_hip.so   hipcc, gfx950: kernels + the C ABI of include/hbdia_hip.h
    hbdia_amd/lib/libhbdia.so       g++ + OpenMP: the C++ API (HBDIA<T>, HBDIAVector<T>,
                                    hbdiaSpMV) and the C handle API of include/hbdia_capi.h,
                                    linked against libhbdia_hip.so
    hbdia_amd/lib/simple_hbdia_example   the reference's single-GPU example, built against
                                    the drop-in headers
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
INCLUDE = ROOT / "include"
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = ROOT / "_build" / "obj"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("HBDIA_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions are disabled for every device compile.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wextra",
             "-Wno-unused-parameter"]

HIP_SOURCES = sorted((CSRC / "hip").glob("*.hip"))
HOST_SOURCES = sorted((CSRC / "host").glob("*.cpp"))


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def _compile_hip(src: Path) -> Path:
    obj = OBJ / (src.stem + ".hip.o")
    _run([HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", "-c", str(src), "-o", str(obj)])
    return obj


SANITIZE = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"]


def _compile_host(src: Path, sanitize: bool = False) -> Path:
    obj = OBJ / (src.stem + (".asan.o" if sanitize else ".o"))
    flags = [*CXX_FLAGS, "-O1", "-g", *SANITIZE] if sanitize else CXX_FLAGS
    _run([CXX, *flags, f"-I{INCLUDE}", "-c", str(src), "-o", str(obj)])
    return obj


def _link_atomic(cmd_prefix: list[str], objs: list[Path], out: Path, extra: list[str]) -> None:
    tmp = out.with_suffix(out.suffix + ".tmp")
    _run([*cmd_prefix, *map(str, objs), "-o", str(tmp), *extra])
    os.replace(tmp, out)


def build_all(jobs: int | None = None, verbose: bool = True, sanitize: bool = False) -> None:
    """Compile every HIP and host source and link the shared libraries."""
    LIB.mkdir(parents=True, exist_ok=True)
    OBJ.mkdir(parents=True, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        hip_objs = list(ex.map(_compile_hip, HIP_SOURCES))
        host_objs = list(ex.map(lambda src: _compile_host(src, sanitize), HOST_SOURCES))
    hip_lib = LIB / "libhbdia_hip.so"
    _link_atomic([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC"], hip_objs, hip_lib, [])
    host_lib = LIB / "libhbdia.so"
    _link_atomic([CXX, "-shared", "-fopenmp", *(SANITIZE if sanitize else [])], host_objs, host_lib,
                 [f"-L{LIB}", "-lhbdia_hip", "-Wl,-rpath,$ORIGIN"])
    example = CSRC / "examples" / "simple_hbdia_example.cpp"
    if example.exists():
        exe = LIB / "simple_hbdia_example"
        _run([CXX, *CXX_FLAGS, f"-I{INCLUDE}", str(example), "-o", str(exe) + ".tmp",
              f"-L{LIB}", "-lhbdia", "-lhbdia_hip", "-Wl,-rpath,$ORIGIN"])
        os.replace(str(exe) + ".tmp", exe)
    if verbose:
        print(f"[hbdia_amd.build] built {hip_lib.name}, {host_lib.name} for {ARCH}", flush=True)


def build_oracle(verbose: bool = True) -> None:
    """Build the CPU oracle (test infrastructure) with its Makefile."""
    make = shutil.which("make")
    if make is None:
        raise RuntimeError("make not found")
    _run([make, "-C", str(ROOT / "oracle"), "-s"])
    if verbose:
        print("[hbdia_amd.build] built oracle/build/libhbdia_oracle.so", flush=True)


if __name__ == "__main__":
    build_all(sanitize="--sanitize" in sys.argv)
    if "--oracle" in sys.argv:
        build_oracle()
