"""ctypes front end of the reference serial CVL, compiled by oracle/Makefile
from nesl/cvl/serial/*.c into oracle/_ref/libcvl_serial.so.

TEST INFRASTRUCTURE ONLY (golden-vector generation, oracle cross-checks,
and the "reference" kind of bench.py's cpu_baseline).  The library exists
only where the reference sources were present at build time; callers test
``available()`` first.

Calls use the CVL names and argument order of nesl/include/cvl.h: numpy
arrays are passed as host ``vec_p`` pointers, Python ints as ``int`` and
Python floats as ``double``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libcvl_serial.so")
ASIS_PATH = os.path.join(HERE, "_ref", "libcvl_serial_asis.so")

_libs: dict[str, C.CDLL] = {}


def available(path: str = LIB_PATH) -> bool:
    return os.path.exists(path)


def lib(path: str = LIB_PATH) -> C.CDLL:
    if path not in _libs:
        _libs[path] = C.CDLL(path)
    return _libs[path]


def _restype(name: str):
    if name.startswith(("ext_vud",)) or (name[4:6] == "ru" and name.endswith("d")):
        return C.c_double
    if name[4:6] == "ru" or name.startswith(("ext_vu", "pk1_luv", "siz_fo")):
        return C.c_int
    if name.endswith(("_scratch",)):
        return C.c_int
    if name.endswith("_inplace"):
        return C.c_uint
    return None


def call(name: str, *args, path: str = LIB_PATH):
    """Call reference function ``name``; arrays become pointers."""
    f = getattr(lib(path), name)
    f.restype = _restype(name)
    cargs = []
    for a in args:
        if isinstance(a, np.ndarray):
            assert a.flags["C_CONTIGUOUS"]
            cargs.append(a.ctypes.data_as(C.c_void_p))
        elif a is None:
            cargs.append(C.c_void_p(None))
        elif isinstance(a, float):
            cargs.append(C.c_double(a))
        else:
            cargs.append(C.c_int(int(a)))
    return f(*cargs)


def scratch_for(units: int) -> np.ndarray:
    """A host scratch area of ``units`` CVL units (8 bytes each)."""
    return np.zeros(max(int(units), 1), dtype=np.float64)
