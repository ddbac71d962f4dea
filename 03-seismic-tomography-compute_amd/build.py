"""Build recipe for the gfx950 library and the C++ drop-in layer.

  lib/libseismic_hip.so  every HIP/C++ source under csrc/ (hipcc, gfx950):
                         the kernels and the C ABI of include/seismic_hip.h
  lib/libseismic.so      cpp/*.cpp (g++, host only): the reference's C++
                         classes (include/seismic/...) over that C ABI
  lib/dropin_driver      tests/cpp/dropin_driver.cpp, the C++ test program

Run directly (`python build.py`) or through __graft_entry__.build().
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = PKG / "_build"
TARGET = LIB / "libseismic_hip.so"
CPP = PKG / "cpp"
HOST_TARGET = LIB / "libseismic.so"
DRIVER_SRC = ROOT / "tests" / "cpp" / "dropin_driver.cpp"
DRIVER = LIB / "dropin_driver"
GXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", f"-I{ROOT / 'include'}",
              f"-I{ROOT / 'include' / 'seismic'}"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("STP_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # hipcc -O3 forms packed FP32 ops by itself; keep them off.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = TARGET.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, TARGET)
    build_host()
    if verbose:
        print(f"built {TARGET.relative_to(ROOT)} from {len(srcs)} sources, {HOST_TARGET.name} and "
              f"{DRIVER.name}", flush=True)
    return TARGET


def _run(cmd: list[str], what: str) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_host() -> None:
    """The C++ drop-in layer (host code only, linked against the HIP library)."""
    srcs = sorted(str(p) for p in CPP.glob("*.cpp"))
    link = [f"-L{LIB}", "-lseismic_hip", "-Wl,-rpath,$ORIGIN"]
    tmp = HOST_TARGET.with_suffix(".so.tmp")
    _run([GXX, *HOST_FLAGS, "-shared", *srcs, *link, "-o", str(tmp)], "host library")
    os.replace(tmp, HOST_TARGET)
    if DRIVER_SRC.exists():
        tmp = DRIVER.with_name(DRIVER.name + ".tmp")
        _run([GXX, *HOST_FLAGS, str(DRIVER_SRC), f"-L{LIB}", "-lseismic", "-lseismic_hip", "-Wl,-rpath,$ORIGIN",
              "-o", str(tmp)], "drop-in driver")
        os.replace(tmp, DRIVER)


if __name__ == "__main__":
    build()
    sys.exit(0)
