"""Build the gfx950 shared library ``varrcwa_amd/libvarrcwa.so`` with hipcc.

Every source in ``csrc/`` is compiled (in parallel) to an object under
``varrcwa_amd/_build/`` and linked into one library.  Run as
``python -m varrcwa_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libvarrcwa.so"
ARCH = os.environ.get("VRC_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          # packed FP32 instructions stay off (they are not used by this FP64 code either)
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-result", "-I", str(CSRC), "-I", str(HERE.parent / "include")]


def sources():
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, min(16, os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cpp_check()
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources (+ {CPP_CHECK.name})", file=sys.stderr)
    return LIB


CPP_CHECK = HERE.parent / "tests" / "cpp" / "cpp_api_check"


def build_cpp_check() -> Path:
    """Host-only C++ program over include/varrcwa.hpp, linked against the library."""
    src = CPP_CHECK.with_suffix(".cpp")
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O2", "-std=c++17", "-I", str(HERE.parent / "include"), str(src), "-o", str(CPP_CHECK),
           "-L", str(HERE), "-lvarrcwa", "-Wl,-rpath,$ORIGIN/../../varrcwa_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return CPP_CHECK


if __name__ == "__main__":
    build()
