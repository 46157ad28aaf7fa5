"""Build recipe of liblgdhip.so (gfx950) and of the C++ host test programs.

Every call compiles from scratch: each HIP source becomes an object with
hipcc --offload-arch=gfx950, then they are linked into lib/liblgdhip.so.
Packed-FP32 instruction selection is disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "lib"
OBJ_DIR = PKG / "_build"
LIB = LIB_DIR / "liblgdhip.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    # no packed FP32 (v_pk_*_f32) instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{REPO / 'include'}",
]

SOURCES = ["runtime.hip", "cellgrid.hip", "jacobi3d.hip", "life2d.hip", "lbm_d3q19.hip", "poisson_batch.hip",
           "unet3d.hip", "learned_guess.hip"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(map(str, cmd))}\n{r.stdout}")
    return r.stdout


def compile_object(src: str) -> Path:
    obj = OBJ_DIR / (Path(src).stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    return obj


def build_library(jobs: int | None = None) -> Path:
    OBJ_DIR.mkdir(parents=True, exist_ok=True)
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_object, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    _run(["make", "-s", "-C", str(REPO / "oracle"), "-B"])
    return REPO / "oracle" / "lib" / "liblgd_oracle.so"


def build_cpp_tests() -> list[Path]:
    """Compile the C++ host programs under tests/cpp (see its Makefile)."""
    mk = REPO / "tests" / "cpp" / "Makefile"
    if not mk.exists():
        return []
    _run(["make", "-s", "-C", str(mk.parent), "-B", f"-j{min(8, os.cpu_count() or 4)}"])
    out = mk.parent / "bin"
    return sorted(out.glob("*")) if out.exists() else []


def build_all() -> None:
    build_library()
    build_oracle()
    build_cpp_tests()


if __name__ == "__main__":
    build_all()
    print("built", LIB, file=sys.stderr)
