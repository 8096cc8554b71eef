"""MI355X-native single-subdomain RHS of milupHPC: keys, sort, octree, Barnes-Hut,
neighbour search and SPH as hand-written HIP kernels for gfx950 behind a C-ABI
(include/mhpc.h).  This Python package binds that C-ABI for tests and benchmarks; the
C++ facade (Miluphpc / ParticleHandler) lives in miluphpc_amd/host/.
"""
from __future__ import annotations

from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libmhpc.so"

__all__ = ["PKG_DIR", "LIB_PATH", "load"]


def load():
    """Import and return the ctypes binding (raises if libmhpc.so is not built)."""
    from . import lib

    return lib
