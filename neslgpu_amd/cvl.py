"""Python mirror of the CVL C API, bound to libcvl_hip.so with ctypes.

Every function of include/cvl.h and include/cvl_hip_ext.h is reachable by
its CVL name with the reference argument order (nesl/include/cvl.h):
``vec_p`` arguments take device tensors (or raw device addresses), ``int`` and
``double`` arguments take Python numbers.  Prototypes are read from the
headers, so the binding cannot drift from the C-ABI.

    from neslgpu_amd import cvl
    lib = cvl.load()
    lib.add_suz(d, s, n, scratch)          # exactly the C call

On top of the raw calls, ``Cvl`` offers allocating helpers used by the tests
and the benchmark (scratch sized by the op's own ``*_scratch`` function,
segment descriptors built by ``mke_fov``).

The library must have been built (python -m neslgpu_amd.build); a missing
library raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
# CVLHIP_LIBRARY names another build of the library (the autotuner's variants)
LIB_PATH = os.environ.get("CVLHIP_LIBRARY") or os.path.join(HERE, "libcvl_hip.so")
HEADERS = [os.path.join(ROOT, "include", "cvl.h"), os.path.join(ROOT, "include", "cvl_hip_ext.h")]

_CTYPES = {
    "void": None, "int": C.c_int, "unsigned int": C.c_uint, "unsigned": C.c_uint,
    "double": C.c_double, "cvl_bool": C.c_int, "vec_p": C.c_void_p, "long long": C.c_longlong,
    "unsigned long long": C.c_ulonglong, "char *": C.c_char_p, "void *": C.c_void_p,
    "float": C.c_float,
}
_PTR_ARG = C.c_void_p   # int *, double *, cvl_bool *, cvl_timer_t *: raw addresses

_PROTO = re.compile(r"^\s*([a-z_][a-z_ ]*?\s*\*?)\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(([^)]*)\)\s*;", re.M)


def parse_prototypes(paths=HEADERS) -> dict:
    """{name: (restype, [argtypes])} for every prototype in the headers."""
    out = {}
    for p in paths:
        text = re.sub(r"/\*.*?\*/", "", open(p).read(), flags=re.S)
        for ret, name, args in _PROTO.findall(text):
            ret = ret.strip()
            rt = _CTYPES.get(ret.replace(" *", " *").strip(), _PTR_ARG if ret.endswith("*") else None)
            ats = []
            a = args.strip()
            if a and a != "void":
                for arg in a.split(","):
                    arg = arg.strip()
                    if "*" in arg and not arg.startswith("vec_p"):
                        ats.append(C.c_char_p if arg.startswith("char") else _PTR_ARG)
                        continue
                    typ = arg.rsplit(" ", 1)[0].strip()
                    ats.append(_CTYPES[typ])
            out[name] = (rt, ats)
    return out


class CvlError(RuntimeError):
    pass


def _addr(x):
    """Device address of a tensor (or pass ints / None through)."""
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):          # numpy array: a host buffer (v2c / c2v)
        return x.ctypes.data
    return x


class Lib:
    """The loaded shared library; attribute access returns checked callables."""

    def __init__(self, path: str = LIB_PATH):
        if not os.path.exists(path):
            raise CvlError(f"{path} is missing: build it with `python -m neslgpu_amd.build`")
        self.path = path
        self.dll = C.CDLL(path)
        self.protos = parse_prototypes()
        for name, (rt, ats) in self.protos.items():
            f = getattr(self.dll, name)
            f.restype = rt
            f.argtypes = ats
        self.dll.cvlhip_set_error_mode(0)
        self._buf = C.create_string_buffer(512)

    def check(self, name="call"):
        code = self.dll.cvlhip_last_error(self._buf, 512)
        if code:
            raise CvlError(f"{name}: {self._buf.value.decode()}")

    def __getattr__(self, name):
        if name.startswith("_") or name not in self.protos:
            raise AttributeError(name)
        f = getattr(self.dll, name)
        rt, ats = self.protos[name]

        def call(*args):
            if len(args) != len(ats):
                raise TypeError(f"{name} takes {len(ats)} arguments, got {len(args)}")
            conv = []
            for a, t in zip(args, ats):
                if t in (C.c_void_p,):
                    conv.append(_addr(a))
                elif t is C.c_double or t is C.c_float:
                    conv.append(float(a))
                elif t is C.c_char_p:
                    conv.append(a)
                else:
                    conv.append(int(a))
            r = f(*conv)
            self.check(name)
            return r

        call.__name__ = name
        return call


_lib: Lib | None = None


def load() -> Lib:
    global _lib
    if _lib is None:
        _lib = Lib()
        import torch
        if torch.cuda.is_available():
            _lib.dll.cvlhip_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream))
            _lib.CVL_init(None)
    return _lib


# ---------------------------------------------------------------- helpers
UNIT = 8   # bytes per CVL unit


class Cvl:
    """Allocating front end over the raw C calls (device tensors in and out)."""

    def __init__(self, device="cuda"):
        import torch
        self.torch = torch
        self.lib = load()
        self.device = torch.device(device)
        self._scr = None

    # -- memory
    def scratch(self, units: int):
        nbytes = max(int(units), 1) * UNIT + 256
        if self._scr is None or self._scr.numel() < nbytes:
            self._scr = self.torch.empty(max(nbytes, 1 << 20), dtype=self.torch.uint8, device=self.device)
        return self._scr

    def empty(self, n, dtype):
        return self.torch.empty(max(int(n), 0), dtype=dtype, device=self.device)

    def tensor(self, a, dtype=None):
        return self.torch.as_tensor(a, dtype=dtype).to(self.device)

    # -- segment descriptors
    def segd(self, lens) -> "Segd":
        torch = self.torch
        lens = self.tensor(lens, torch.int32).contiguous()
        m = lens.numel()
        n = int(lens.to(torch.int64).sum().item()) if m else 0
        units = self.lib.siz_fos(n, m)
        sd = self.torch.zeros(units * UNIT // 4 + 64, dtype=torch.int32, device=self.device)
        scr = self.scratch(self.lib.mke_fov_scratch(n, m))
        self.lib.mke_fov(sd, lens, n, m, scr)
        return Segd(sd, n, m)


class Segd:
    """A device segment descriptor with its element count n and segment count m."""

    def __init__(self, buf, n, m):
        self.buf, self.n, self.m = buf, n, m

    def data_ptr(self):
        return self.buf.data_ptr()
