"""libgeodecomp_amd -- MI355X-native HipSimulator sweeps for LibGeoDecomp cells.

The product is the C-ABI library `lib/liblgdhip.so` (gfx950 HIP kernels,
declared in include/lgdhip.h) plus the header-only C++ HipSimulator under
`src/libgeodecomp/`.  This Python package only locates, builds and binds
them (see `lgdhip.py`, `build.py`).
"""
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
INCLUDE_DIR = REPO_DIR / "include"
CXX_INCLUDE_DIR = PKG_DIR / "src"

__all__ = ["PKG_DIR", "REPO_DIR", "INCLUDE_DIR", "CXX_INCLUDE_DIR"]
