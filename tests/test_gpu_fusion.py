"""Generated fused kernels against the oracle, one super-op at a time.

Every super-op of the reference programs' definitions (and of this
package's SpMV program) is evaluated on random inputs by its fused HIP
kernel (called through the fused library's C entry point) and by the oracle
applying the same expression tree one CVL elementwise operation at a time.
Integer, boolean and + - * / results must match bit for bit; results that
pass through a libm transcendental (EXP, LOG, ATAN, ...) within 4 ulp, since
device and host libm may round differently.

A super-op whose root is a segmented reduction must give exactly the bits
of the CVL reduction (add_red, max_red, ...) applied to the expression's
values, since it runs the same kernel with the same tiles; max/min also
equal the oracle's fold, and + is within the compensated-sum bound.
"""
import ctypes as C
import glob
import os

import numpy as np
import pytest

from gpu_util import host
from gpu_util import fp_add_tol_vec
from neslgpu_amd.build import FUSED_DIR, PROGRAMS_DIR, SUPEROPS_DIR
from neslgpu_amd.fusion import superops as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEFS = sorted(glob.glob(os.path.join(SUPEROPS_DIR, "*.superops")) + glob.glob(os.path.join(PROGRAMS_DIR, "*.superops")))
TY = {"INT": "z", "FLOAT": "d", "BOOL": "b"}
BIN = {"+": "add", "-": "sub", "*": "mul", "/": "div", "%": "mod", "<": "les", "<=": "leq", ">": "grt",
       ">=": "geq", "=": "eql", "!=": "neq", "LSHIFT": "lsh", "RSHIFT": "rsh", "AND": "and", "OR": "ior",
       "XOR": "xor", "MAX": "max", "MIN": "min"}
UN = {("NOT", "BOOL"): "notb", ("NOT", "INT"): "notz", "B_TO_I": "int_b", "I_TO_F": "dbl_z",
      "I_TO_B": "boo_z", "ROUND": "rou", "FLOOR": "flr", "CEIL": "cei", "TRUNC": "trn", "LOG": "log",
      "SQRT": "sqt", "EXP": "exp", "SIN": "sin", "COS": "cos", "TAN": "tan", "ASIN": "asn", "ACOS": "acs",
      "ATAN": "atn", "SINH": "snh", "COSH": "csh", "TANH": "tnh"}
TRANSCENDENTAL = {"LOG", "EXP", "SIN", "COS", "TAN", "ASIN", "ACOS", "ATAN", "SINH", "COSH", "TANH"}


def oracle_eval(e, args, n):
    if isinstance(e, S.Arg):
        return args[e.index]
    if isinstance(e, S.Const):
        if e.type == "FLOAT":
            return np.full(n, float(e.value))
        return np.full(n, 1 if e.value == "T" else (0 if e.value == "F" else int(e.value)), np.int32)
    if e.op == "DIST":
        return oracle_eval(e.args[0], args, n)
    if e.op == "GATHER":
        src = args[e.args[0].index]
        idx = oracle_eval(e.args[1], args, n).astype(np.int64)
        ok = (idx >= 0) & (idx < src.size)
        zero = src.dtype.type(0)
        return np.where(ok, src[np.clip(idx, 0, max(src.size - 1, 0))] if src.size else zero, zero)
    if e.op == "SELECT":
        f, a, b = (oracle_eval(x, args, n) for x in e.args)
        return O.sel("d" if a.dtype == np.float64 else "z", f, a, b)
    vals = [oracle_eval(x, args, n) for x in e.args]
    operand = e.type or S._OPERAND[e.op]
    if e.op in BIN:
        return O.ew2(BIN[e.op], TY[operand], vals[0], vals[1])
    key = (e.op, operand) if (e.op, operand) in UN else e.op
    return O.ew1(UN[key], vals[0])


def has_transcendental(e):
    return isinstance(e, S.Apply) and (e.op in TRANSCENDENTAL or any(has_transcendental(x) for x in e.args))


def random_arg(t, n, rng):
    if t == "INT":
        return rng.integers(1, 60, n).astype(np.int32)          # non-zero: / and % stay defined
    if t == "BOOL":
        return rng.integers(0, 2, n).astype(np.int32)
    return rng.uniform(0.25, 4.0, n)


def check_reduce(cv, o, vals, lens, got):
    """A fused segmented reduction against the CVL reduction of the same
    values (same bits) and against the oracle."""
    import torch
    rop = {"+_REDUCE": "add", "MAX_REDUCE": "max", "MIN_REDUCE": "min"}[o.reduce[0]]
    ty = "d" if o.result == "FLOAT" else "z"
    n, m = vals.size, lens.size
    sd = cv.segd(lens)
    dv = torch.from_numpy(np.ascontiguousarray(vals)).to(cv.device)
    ref = torch.empty(m, dtype=dv.dtype, device=cv.device)
    name = f"{rop}_re{ty}"
    getattr(cv.lib, name)(ref, dv, sd, n, m, cv.scratch(getattr(cv.lib, name + "_scratch")(n, m)))
    r = host(ref)
    if ty == "d":
        assert np.array_equal(got.view(np.int64), r.view(np.int64)), o.text
    else:
        assert np.array_equal(got, r), o.text
    if ty == "d" and rop == "add":
        ref_s, absum = O.segreduce_add_neumaier(vals, lens)
        assert np.all(np.abs(got - ref_s) <= fp_add_tol_vec(absum, lens)), o.text
    else:
        assert np.array_equal(got, O.segreduce(rop, ty, vals, lens)), o.text


@pytest.fixture(scope="module")
def cv():
    from neslgpu_amd.cvl import Cvl
    return Cvl()


@pytest.mark.parametrize("path", DEFS, ids=os.path.basename)
def test_superops_match_oracle(cv, path):
    import torch
    name = os.path.basename(path)[: -len(".superops")]
    lib = C.CDLL(os.path.join(FUSED_DIR, name + ".so"))
    ops = S.parse(open(path).read())
    rng = np.random.default_rng(5)
    for n in (1, 1000, 100_003):
        for o in ops:
            print("superop", name, o.number, "n", n, flush=True)
            gsrc = S.gather_sources(o.body)
            gseg = {v: k for k, v in gsrc.items()}
            hargs, dargs, rlens = [], [], None
            for i, t in enumerate(o.params):
                if i in gsrc:                          # a gathered vector: its own length
                    a = random_arg(t, 37 + n // 3, rng)
                elif t == "SEGDES" and i in gseg:      # describes the gathered vector
                    hargs.append(None)
                    dargs.append(cv.segd(np.array([hargs[gseg[i]].size], np.int32)))
                    continue
                elif t == "SEGDES" and o.reduce and i == o.reduce[1]:
                    rlens = np.diff(np.concatenate([[0], np.sort(rng.integers(0, n + 1, max(1, n // 5))), [n]]))
                    rlens = rlens.astype(np.int32)
                    hargs.append(None)
                    dargs.append(cv.segd(rlens))
                    continue
                elif t == "SEGDES":                    # only its length matters
                    hargs.append(None)
                    dargs.append(cv.segd(np.array([n], np.int32)))
                    continue
                else:
                    a = random_arg(t, n, rng)
                hargs.append(a)
                dargs.append(torch.from_numpy(a).to(cv.device))
            want = oracle_eval(o.body, hargs, n)
            rows = rlens.size if o.reduce else n
            out = torch.empty(rows, dtype=torch.float64 if o.result == "FLOAT" else torch.int32, device=cv.device)
            fn = getattr(lib, f"fused{o.number}")
            fn.restype = None
            fn.argtypes = [C.c_void_p] * (1 + len(dargs)) + [C.c_int, C.c_void_p]
            scr = None
            if o.reduce:
                sfn = getattr(lib, f"fused{o.number}_scratch")
                sfn.restype = C.c_int
                scr = cv.scratch(sfn(n, rlens.size))
            fn(out.data_ptr(), *[d.data_ptr() for d in dargs], n, None if scr is None else scr.data_ptr())
            got = host(out)
            if o.reduce:
                check_reduce(cv, o, want, rlens, got)
                continue
            if o.result == "FLOAT" and has_transcendental(o.body):
                assert np.allclose(got, want, rtol=4 * 2.0 ** -52, atol=0, equal_nan=True), o.text
            elif o.result == "FLOAT":
                assert np.array_equal(got.view(np.int64), want.view(np.int64)), o.text
            else:
                assert np.array_equal(got, want), o.text
