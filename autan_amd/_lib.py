"""Loading of the native libraries built in-tree by autan_amd/build.py.

The HIP library is required: nothing falls back to a CPU path when it is missing.
"""

from __future__ import annotations

import ctypes
import functools
from pathlib import Path

HERE = Path(__file__).resolve().parent
MESH_LIB = HERE / "libmxmesh.so"
EULER_LIB = HERE / "libmxeuler.so"


@functools.cache
def mesh() -> ctypes.CDLL:
    if not MESH_LIB.exists():
        raise RuntimeError(f"{MESH_LIB} is not built: run autan_amd/build.py (or __graft_entry__.build())")
    lib = ctypes.CDLL(str(MESH_LIB))
    c = ctypes
    lib.mxmesh_destroy.argtypes = [c.c_void_p]
    lib.mxmesh_destroy.restype = None
    lib.mxmesh_sizes.argtypes = [c.c_void_p, c.POINTER(c.c_int64)]
    lib.mxmesh_sizes.restype = None
    return lib


@functools.cache
def euler() -> ctypes.CDLL:
    if not EULER_LIB.exists():
        raise RuntimeError(f"{EULER_LIB} is not built: run autan_amd/build.py (or __graft_entry__.build())")
    return ctypes.CDLL(str(EULER_LIB))
