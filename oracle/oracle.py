"""ctypes front end of the CPU oracle (oracle/cvl_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library
(neslgpu_amd) never imports this module.

All functions take and return numpy arrays.  int32 / cvl_bool vectors are
``np.int32``; doubles are ``np.float64``.  Segment descriptors are given as
the vector of segment lengths, which is what the serial CVL's descriptor is
(ref nesl/cvl/serial/facilt.c:106-123).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# Operator and type codes (cvl_oracle.c enums).
ADD, MUL, MIN, MAX, AND, IOR, XOR = range(7)
Z, D, B = range(3)
OPS = {"add": ADD, "mul": MUL, "min": MIN, "max": MAX, "and": AND, "ior": IOR, "xor": XOR}
TYPES = {"z": Z, "d": D, "b": B}
EW2 = {n: k for k, n in enumerate(
    ["max", "min", "add", "sub", "mul", "div", "grt", "les", "geq", "leq", "eql", "neq",
     "lsh", "rsh", "mod", "and", "ior", "xor"])}
EW1 = {n: k for k, n in enumerate(
    ["notb", "notz", "int_d", "int_b", "dbl_z", "boo_z", "flr", "cei", "trn", "rou", "exp",
     "log", "sqt", "sin", "cos", "tan", "asn", "acs", "atn", "snh", "csh", "tnh"])}

_lib = None


def build() -> None:
    """Compile liboracle.so with gcc (fast; used when the .so is missing)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = C.c_void_p
I64 = C.c_int64
I32 = C.c_int32


def _declare(L: C.CDLL) -> None:
    sig = {
        "orc_scan": (C.c_int, [C.c_int, C.c_int, P, P, I64]),
        "orc_segscan": (C.c_int, [C.c_int, C.c_int, P, P, P, I64]),
        "orc_reduce": (C.c_int, [C.c_int, C.c_int, P, P, I64]),
        "orc_segreduce": (C.c_int, [C.c_int, C.c_int, P, P, P, I64]),
        "orc_segscan_add_neumaier": (None, [P, P, P, P, I64]),
        "orc_segreduce_add_neumaier": (None, [P, P, P, P, I64]),
        "orc_offsets": (None, [P, P, I64]),
        "orc_ind_luz": (None, [P, I32, I32, I64]),
        "orc_ind_lez": (None, [P, P, P, P, I64]),
        "orc_dis_ve": (None, [C.c_int, P, P, P, I64]),
        "orc_ext_ve": (None, [C.c_int, P, P, P, P, I64]),
        "orc_rep_ve": (None, [C.c_int, P, P, P, P, I64]),
        "orc_pk1_luv": (I64, [P, I64]),
        "orc_pk2": (I64, [C.c_int, P, P, P, I64]),
        "orc_pk1_lev": (None, [P, P, P, I64]),
        "orc_smp_pu": (None, [C.c_int, P, P, P, I64]),
        "orc_smp_pe": (None, [C.c_int, P, P, P, P, I64]),
        "orc_bck_pu": (None, [C.c_int, P, P, P, I64]),
        "orc_bck_pe": (None, [C.c_int, P, P, P, P, I64, P, I64, I64]),
        "orc_fpm_pu": (None, [C.c_int, P, P, P, P, I64]),
        "orc_fpm_pe": (None, [C.c_int, P, P, P, P, P, I64, P, I64, I64]),
        "orc_bfp_pu": (None, [C.c_int, P, P, P, P, I64]),
        "orc_bfp_pe": (None, [C.c_int, P, P, P, P, P, I64, P, I64, I64]),
        "orc_dpe_pu": (None, [C.c_int, P, P, P, P, I64, I64]),
        "orc_dpe_pe": (None, [C.c_int, P, P, P, P, P, I64, P, I64, I64, I64]),
        "orc_dfp_pu": (None, [C.c_int, P, P, P, P, P, I64, I64]),
        "orc_dfp_pe": (None, [C.c_int, P, P, P, P, P, P, I64, P, I64, I64, I64]),
        "orc_rank_z": (None, [P, P, I64, P, I64, C.c_int, P]),
        "orc_rank_d": (None, [P, P, I64, P, I64, C.c_int, P]),
        "orc_ew2_z": (C.c_int, [C.c_int, P, P, P, I64]),
        "orc_ew2_d": (C.c_int, [C.c_int, P, P, P, P, I64]),
        "orc_ew2_b": (C.c_int, [C.c_int, P, P, P, I64]),
        "orc_ew1": (C.c_int, [C.c_int, P, P, I64]),
        "orc_sel": (None, [C.c_int, P, P, P, P, I64]),
        "orc_lrand48_next": (I32, [P]),
        "orc_rnd_wuz": (None, [P, P, I64, P]),
        "orc_splitmix_u32": (None, [P, I64, C.c_uint64]),
        "orc_splitmix_f64": (None, [P, I64, C.c_uint64, C.c_double, C.c_double]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _ty(a: np.ndarray, ty: str | None = None) -> int:
    if ty is not None:
        return TYPES[ty]
    return D if a.dtype == np.float64 else Z


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _dt(ty: str):
    return np.float64 if ty == "d" else np.int32


# ---------------------------------------------------------------- scans
def scan(op: str, ty: str, s) -> np.ndarray:
    s = _c(s, _dt(ty))
    d = np.empty_like(s)
    rc = lib().orc_scan(OPS[op], TYPES[ty], _p(d), _p(s), s.size)
    assert rc == 0, (op, ty)
    return d


def segscan(op: str, ty: str, s, lens) -> np.ndarray:
    s = _c(s, _dt(ty)); lens = _c(lens, np.int32)
    assert int(lens.astype(np.int64).sum()) == s.size
    d = np.empty_like(s)
    rc = lib().orc_segscan(OPS[op], TYPES[ty], _p(d), _p(s), _p(lens), lens.size)
    assert rc == 0, (op, ty)
    return d


def reduce(op: str, ty: str, s):
    s = _c(s, _dt(ty))
    out = np.zeros(1, dtype=_dt(ty))
    rc = lib().orc_reduce(OPS[op], TYPES[ty], _p(out), _p(s), s.size)
    assert rc == 0
    return out[0]


def segreduce(op: str, ty: str, s, lens) -> np.ndarray:
    s = _c(s, _dt(ty)); lens = _c(lens, np.int32)
    d = np.empty(lens.size, dtype=_dt(ty))
    rc = lib().orc_segreduce(OPS[op], TYPES[ty], _p(d), _p(s), _p(lens), lens.size)
    assert rc == 0
    return d


def segscan_add_neumaier(s, lens):
    s = _c(s, np.float64); lens = _c(lens, np.int32)
    out = np.empty_like(s); absp = np.empty_like(s)
    lib().orc_segscan_add_neumaier(_p(out), _p(absp), _p(s), _p(lens), lens.size)
    return out, absp


def segreduce_add_neumaier(s, lens):
    s = _c(s, np.float64); lens = _c(lens, np.int32)
    out = np.empty(lens.size); absum = np.empty(lens.size)
    lib().orc_segreduce_add_neumaier(_p(out), _p(absum), _p(s), _p(lens), lens.size)
    return out, absum


def offsets(lens) -> np.ndarray:
    lens = _c(lens, np.int32)
    o = np.empty(lens.size + 1, dtype=np.int64)
    lib().orc_offsets(_p(o), _p(lens), lens.size)
    return o


# ------------------------------------------------- index/dist/ext/rep
def ind_luz(init: int, stride: int, n: int) -> np.ndarray:
    d = np.empty(n, dtype=np.int32)
    lib().orc_ind_luz(_p(d), init, stride, n)
    return d


def ind_lez(init, stride, lens) -> np.ndarray:
    init = _c(init, np.int32); stride = _c(stride, np.int32); lens = _c(lens, np.int32)
    d = np.empty(int(lens.sum()), dtype=np.int32)
    lib().orc_ind_lez(_p(d), _p(init), _p(stride), _p(lens), lens.size)
    return d


def dis_ve(ty: str, v, lens) -> np.ndarray:
    v = _c(v, _dt(ty)); lens = _c(lens, np.int32)
    d = np.empty(int(lens.sum()), dtype=_dt(ty))
    lib().orc_dis_ve(TYPES[ty], _p(d), _p(v), _p(lens), lens.size)
    return d


def ext_ve(ty: str, s, idx, lens) -> np.ndarray:
    s = _c(s, _dt(ty)); idx = _c(idx, np.int32); lens = _c(lens, np.int32)
    d = np.empty(lens.size, dtype=_dt(ty))
    lib().orc_ext_ve(TYPES[ty], _p(d), _p(s), _p(idx), _p(lens), lens.size)
    return d


def rep_ve(ty: str, d, idx, v, lens) -> np.ndarray:
    d = _c(d, _dt(ty)).copy(); idx = _c(idx, np.int32); v = _c(v, _dt(ty))
    lens = _c(lens, np.int32)
    lib().orc_rep_ve(TYPES[ty], _p(d), _p(idx), _p(v), _p(lens), lens.size)
    return d


# ------------------------------------------------------------ pack
def pk1_luv(f) -> int:
    f = _c(f, np.int32)
    return int(lib().orc_pk1_luv(_p(f), f.size))


def pk2(ty: str, s, f) -> np.ndarray:
    s = _c(s, _dt(ty)); f = _c(f, np.int32)
    d = np.empty(pk1_luv(f), dtype=_dt(ty))
    lib().orc_pk2(TYPES[ty], _p(d), _p(s), _p(f), s.size)
    return d


def pk1_lev(f, lens) -> np.ndarray:
    f = _c(f, np.int32); lens = _c(lens, np.int32)
    ds = np.empty(lens.size, dtype=np.int32)
    lib().orc_pk1_lev(_p(ds), _p(f), _p(lens), lens.size)
    return ds


# ---------------------------------------------------------- permutes
def smp_pu(ty, s, idx, n_d=None):
    s = _c(s, _dt(ty)); idx = _c(idx, np.int32)
    d = np.zeros(s.size if n_d is None else n_d, dtype=_dt(ty))
    lib().orc_smp_pu(TYPES[ty], _p(d), _p(s), _p(idx), s.size)
    return d


def smp_pe(ty, s, idx, lens):
    s = _c(s, _dt(ty)); idx = _c(idx, np.int32); lens = _c(lens, np.int32)
    d = np.zeros_like(s)
    lib().orc_smp_pe(TYPES[ty], _p(d), _p(s), _p(idx), _p(lens), lens.size)
    return d


def bck_pu(ty, s, idx):
    s = _c(s, _dt(ty)); idx = _c(idx, np.int32)
    d = np.empty(idx.size, dtype=_dt(ty))
    lib().orc_bck_pu(TYPES[ty], _p(d), _p(s), _p(idx), idx.size)
    return d


def bck_pe(ty, s, idx, lens_s, lens_d):
    s = _c(s, _dt(ty)); idx = _c(idx, np.int32)
    lens_s = _c(lens_s, np.int32); lens_d = _c(lens_d, np.int32)
    d = np.empty(idx.size, dtype=_dt(ty))
    lib().orc_bck_pe(TYPES[ty], _p(d), _p(s), _p(idx), _p(lens_s), lens_s.size,
                     _p(lens_d), lens_d.size, idx.size)
    return d


def fpm_pu(ty, s, idx, f, n_d):
    s = _c(s, _dt(ty)); idx = _c(idx, np.int32); f = _c(f, np.int32)
    d = np.zeros(n_d, dtype=_dt(ty))
    lib().orc_fpm_pu(TYPES[ty], _p(d), _p(s), _p(idx), _p(f), s.size)
    return d


def fpm_pe(ty, s, idx, f, lens_s, lens_d, d0=None):
    s = _c(s, _dt(ty)); idx = _c(idx, np.int32); f = _c(f, np.int32)
    lens_s = _c(lens_s, np.int32); lens_d = _c(lens_d, np.int32)
    n_d = int(lens_d.sum())
    d = np.zeros(n_d, dtype=_dt(ty)) if d0 is None else _c(d0, _dt(ty)).copy()
    lib().orc_fpm_pe(TYPES[ty], _p(d), _p(s), _p(idx), _p(f), _p(lens_s), lens_s.size,
                     _p(lens_d), lens_d.size, s.size)
    return d


def bfp_pu(ty, s, idx, f):
    s = _c(s, _dt(ty)); idx = _c(idx, np.int32); f = _c(f, np.int32)
    d = np.empty(idx.size, dtype=_dt(ty))
    lib().orc_bfp_pu(TYPES[ty], _p(d), _p(s), _p(idx), _p(f), idx.size)
    return d


def bfp_pe(ty, s, idx, f, lens_s, lens_d):
    s = _c(s, _dt(ty)); idx = _c(idx, np.int32); f = _c(f, np.int32)
    lens_s = _c(lens_s, np.int32); lens_d = _c(lens_d, np.int32)
    d = np.empty(idx.size, dtype=_dt(ty))
    lib().orc_bfp_pe(TYPES[ty], _p(d), _p(s), _p(idx), _p(f), _p(lens_s), lens_s.size,
                     _p(lens_d), lens_d.size, idx.size)
    return d


def dpe_pu(ty, s, idx, v):
    s = _c(s, _dt(ty)); idx = _c(idx, np.int32); v = _c(v, _dt(ty))
    d = np.empty_like(v)
    lib().orc_dpe_pu(TYPES[ty], _p(d), _p(s), _p(idx), _p(v), s.size, v.size)
    return d


def dpe_pe(ty, s, idx, v, lens_s, lens_d):
    s = _c(s, _dt(ty)); idx = _c(idx, np.int32); v = _c(v, _dt(ty))
    lens_s = _c(lens_s, np.int32); lens_d = _c(lens_d, np.int32)
    d = np.empty_like(v)
    lib().orc_dpe_pe(TYPES[ty], _p(d), _p(s), _p(idx), _p(v), _p(lens_s), lens_s.size,
                     _p(lens_d), lens_d.size, s.size, v.size)
    return d


def dfp_pu(ty, s, idx, f, v):
    s = _c(s, _dt(ty)); idx = _c(idx, np.int32); f = _c(f, np.int32); v = _c(v, _dt(ty))
    d = np.empty_like(v)
    lib().orc_dfp_pu(TYPES[ty], _p(d), _p(s), _p(idx), _p(f), _p(v), s.size, v.size)
    return d


def dfp_pe(ty, s, idx, f, v, lens_s, lens_d):
    s = _c(s, _dt(ty)); idx = _c(idx, np.int32); f = _c(f, np.int32); v = _c(v, _dt(ty))
    lens_s = _c(lens_s, np.int32); lens_d = _c(lens_d, np.int32)
    d = np.empty_like(v)
    lib().orc_dfp_pe(TYPES[ty], _p(d), _p(s), _p(idx), _p(f), _p(v), _p(lens_s),
                     lens_s.size, _p(lens_d), lens_d.size, s.size, v.size)
    return d


# ------------------------------------------------------------- rank
def rank(s, up: bool = True, lens=None) -> np.ndarray:
    """Stable rank of int32 or float64 keys (segmented when lens given)."""
    s = np.ascontiguousarray(s)
    n = s.size
    work = np.empty(3 * max(n, 1), dtype=np.int32)
    d = np.empty(n, dtype=np.int32)
    lp = None
    m = 0
    if lens is not None:
        lens = _c(lens, np.int32); lp = _p(lens); m = lens.size
    if s.dtype == np.float64:
        lib().orc_rank_d(_p(d), _p(s), n, lp, m, int(up), _p(work))
    else:
        s = _c(s, np.int32)
        lib().orc_rank_z(_p(d), _p(s), n, lp, m, int(up), _p(work))
    return d


# ------------------------------------------------------- elementwise
def ew2(op: str, ty: str, a, b) -> np.ndarray:
    a = _c(a, _dt(ty)); b = _c(b, _dt(ty))
    n = a.size
    if ty == "d":
        d64 = np.empty(n, np.float64); d32 = np.empty(n, np.int32)
        rc = lib().orc_ew2_d(EW2[op], _p(d64), _p(d32), _p(a), _p(b), n)
        assert rc == 0, op
        return d32 if op in ("grt", "les", "geq", "leq", "eql", "neq") else d64
    d = np.empty(n, np.int32)
    f = lib().orc_ew2_z if ty == "z" else lib().orc_ew2_b
    rc = f(EW2[op], _p(d), _p(a), _p(b), n)
    assert rc == 0, op
    return d


_EW1_IN_D = {"int_d", "flr", "cei", "trn", "rou", "exp", "log", "sqt", "sin", "cos", "tan",
             "asn", "acs", "atn", "snh", "csh", "tnh"}
_EW1_OUT_D = {"dbl_z", "exp", "log", "sqt", "sin", "cos", "tan", "asn", "acs", "atn",
              "snh", "csh", "tnh"}


def ew1(op: str, s) -> np.ndarray:
    s = _c(s, np.float64 if op in _EW1_IN_D else np.int32)
    d = np.empty(s.size, np.float64 if op in _EW1_OUT_D else np.int32)
    rc = lib().orc_ew1(EW1[op], _p(d), _p(s), s.size)
    assert rc == 0
    return d


def sel(ty, f, a, b):
    f = _c(f, np.int32); a = _c(a, _dt(ty)); b = _c(b, _dt(ty))
    d = np.empty_like(a)
    lib().orc_sel(TYPES[ty], _p(d), _p(f), _p(a), _p(b), a.size)
    return d


# -------------------------------------------------------------- RNG
class Lrand48:
    """The 48-bit state behind lrand48 (glibc starts an unseeded process at 0)."""

    def __init__(self, state: int = 0):
        self.state = C.c_uint64(state)

    @staticmethod
    def seeded(seed: int) -> "Lrand48":
        """State after srand48(seed): high 32 bits = seed, low 16 = 0x330E."""
        return Lrand48(((seed & 0xFFFFFFFF) << 16) | 0x330E)

    def next(self) -> int:
        return int(lib().orc_lrand48_next(C.byref(self.state)))

    def rnd_wuz(self, s) -> np.ndarray:
        s = _c(s, np.int32)
        d = np.empty_like(s)
        lib().orc_rnd_wuz(_p(d), _p(s), s.size, C.byref(self.state))
        return d


def splitmix_u32(n: int, seed: int) -> np.ndarray:
    d = np.empty(n, dtype=np.uint32)
    lib().orc_splitmix_u32(_p(d), n, seed)
    return d


def splitmix_f64(n: int, seed: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    d = np.empty(n, dtype=np.float64)
    lib().orc_splitmix_f64(_p(d), n, seed, lo, hi)
    return d
