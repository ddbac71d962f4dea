"""Build recipe for the CPU oracle (test infrastructure only).

  liboracle.so  — the FP64 C restatement in oracle/*.c (gcc, OpenMP)
  liboracle_timing.so — the same sources with the reference's release flags,
                  for the CPU baseline timings only
  _ref/         — the reference's own C++ sources that compile as they are,
                  built from /root/reference when it is present (never copied)

Flags: -O3 -march=x86-64-v3 (AVX2+FMA, runs on any current x86 host; the
reference builds with -march=native -mavx2 -mfma, CMakeLists.txt:25-33) and
no -ffast-math, which would remove the finite checks the oracle restates.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REF = Path(os.environ.get("STP_REFERENCE", "/root/reference"))
LIB = HERE / "liboracle.so"
REF_OUT = HERE / "_ref"

CFLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-Wall", "-Wno-unknown-pragmas"]
# timing-only copy for the CPU baseline: the reference's release flags
# (CMakeLists.txt:25, 31, 33: -O3 -march=native -ffast-math -mavx2 -mfma) with
# -march=x86-64-v3 for native, since it is built here and timed on another
# host; -ffast-math drops the finite checks, so it is never the checker
TIMING_LIB = HERE / "liboracle_timing.so"
TIMING_FLAGS = ["-O3", "-march=x86-64-v3", "-mavx2", "-mfma", "-ffast-math", "-fopenmp", "-fPIC", "-shared",
                "-std=c11", "-Wno-unknown-pragmas"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_oracle() -> Path:
    srcs = sorted(str(p) for p in HERE.glob("*.c"))
    for out, flags in ((LIB, CFLAGS), (TIMING_LIB, TIMING_FLAGS)):
        tmp = out.with_suffix(".so.tmp")
        _run(["gcc", *flags, *srcs, "-lm", "-o", str(tmp)])
        os.replace(tmp, out)
    return LIB


def build_ref() -> list[Path]:
    """Reference translation units that build from their own sources, each
    linked with a small driver of ours (ref_shim.cpp) exposing extern "C"
    entry points.  Skipped when the reference tree is absent."""
    if not (REF / "src" / "cpp").is_dir():
        return []
    REF_OUT.mkdir(exist_ok=True)
    cpp = REF / "src" / "cpp"
    out = REF_OUT / "libref.so"
    tmp = out.with_suffix(".so.tmp")
    srcs = [cpp / "sparse" / "csr_matrix.cpp", cpp / "rtm" / "rtm_engine.cpp", cpp / "core" / "config.cpp",
            cpp / "core" / "grid.cpp", cpp / "solver" / "boundary.cpp", HERE / "ref_shim.cpp"]
    if not all(s.exists() for s in srcs):
        return []
    # rtm_engine.h uses std::function without including <functional>
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-include", "functional", f"-I{cpp}",
          *map(str, srcs), "-o", str(tmp)])
    os.replace(tmp, out)
    return [out]


def build(verbose: bool = True) -> None:
    p = build_oracle()
    refs = build_ref() if (HERE / "ref_shim.cpp").exists() else []
    if verbose:
        print(f"built {p.name}" + (f" and {', '.join(r.name for r in refs)}" if refs else ""), flush=True)


if __name__ == "__main__":
    build()
    sys.exit(0)
