"""Fusion back end without a GPU: super-op parsing and typing, generated
source, and the descriptor tables of the built fused libraries (read with
ctypes; nothing is launched)."""
import ctypes as C
import glob
import os

import pytest

from neslgpu_amd.build import FUSED_DIR, PROGRAMS_DIR, SUPEROPS_DIR
from neslgpu_amd.fusion import codegen, superops as S

DEFS = sorted(glob.glob(os.path.join(SUPEROPS_DIR, "*.superops")) + glob.glob(os.path.join(PROGRAMS_DIR, "*.superops")))


def test_parse_and_type():
    op = S.parse_line("fused OP3#2 ($0 : INT, $1 : SEGDES) = (/ INT @ $0 (DIST INT @ (CONST INT 2) $1))")
    assert (op.number, op.uses, op.params, op.result) == (3, 2, ["INT", "SEGDES"], "INT")
    op = S.parse_line("fused OP1#1 ($0 : INT, $1 : SEGDES) = (< INT @ $0 (DIST INT @ (CONST INT 0) $1))")
    assert op.result == "BOOL"
    op = S.parse_line("fused OP2#1 ($0 : INT) = (I_TO_F @ $0)")
    assert op.result == "FLOAT"
    op = S.parse_line("fused OP0#1 ($0 : BOOL, $1 : FLOAT, $2 : FLOAT) = (SELECT FLOAT @ $0 $1 $2)")
    assert op.result == "FLOAT"
    op = S.parse_line("fused OP0#1 ($0 : INT, $1 : FLOAT, $2 : SEGDES) = (GATHER FLOAT @ $1 $0 $2)")
    assert op.result == "FLOAT" and S.gather_sources(op.body) == {1: 2}
    op = S.parse_line("fused OP1#1 ($0 : FLOAT, $1 : SEGDES) = (+_REDUCE FLOAT @ (* FLOAT @ $0 $0) $1)")
    assert op.result == "FLOAT" and op.reduce == ("+_REDUCE", 1)


@pytest.mark.parametrize("bad", [
    "fused OP0#1 ($0 : INT) = (+ INT @ $0 $1)",                      # unknown parameter
    "fused OP0#1 ($0 : INT, $1 : FLOAT) = (+ INT @ $0 $1)",          # operand type
    "fused OP0#1 ($0 : SEGDES) = (+ INT @ $0 $0)",                   # segdes as a value
    "fused OP0#1 ($0 : INT, $1 : SEGDES) = (DIST INT @ $0 $1)",      # DIST of a vector
    "fused OP0#1 ($0 : INT) = (RAND @ $0)",                          # not fusable
    "fused OP0#1 ($0 : INT) = (+ INT @ $0",                          # unbalanced
    "fused OP0#1 ($0 : INT, $1 : SEGDES) = (+ INT @ (+_REDUCE INT @ $0 $1) $0)",   # reduce not at the root
    "fused OP0#1 ($0 : BOOL, $1 : SEGDES) = (+_REDUCE BOOL @ $0 $1)",              # reduce of bools
])
def test_rejects_bad_definitions(bad):
    with pytest.raises((S.DefinitionError, IndexError)):
        S.parse_line(bad)


@pytest.mark.parametrize("path", DEFS, ids=os.path.basename)
def test_reference_definitions_generate(path):
    ops = S.parse(open(path).read())
    src = codegen.generate(ops, "x")
    for o in ops:
        assert f'extern "C" void fused{o.number}(' in src
        if o.reduce:
            assert f"struct InFused{o.number} {{" in src       # loader for the shared reduce kernel
        else:
            assert f"__global__ void __launch_bounds__(256) k_fused{o.number}(" in src
    assert src.count('{VCODE_FUSED_TOKEN, "fused') == len(ops)


class VopDes(C.Structure):
    _fields_ = [("vop", C.c_int), ("vopname", C.c_char_p), ("arg_num", C.c_int), ("result_num", C.c_int),
                ("arg_type", C.c_int * 40), ("arg_len", C.c_uint * 40), ("result_type", C.c_int * 3),
                ("result_len", C.c_uint * 3), ("arg_for_scratch", C.c_int * 40), ("cvl_desc", C.c_int)]


TYPE = {"BOOL": 3, "INT": 4, "FLOAT": 5, "SEGDES": 6}


@pytest.mark.parametrize("path", DEFS, ids=os.path.basename)
def test_built_library_tables(path):
    name = os.path.basename(path)[: -len(".superops")]
    so = os.path.join(FUSED_DIR, name + ".so")
    if not os.path.exists(so):
        pytest.skip("fused libraries not built")
    lib = C.CDLL(so)
    ops = S.parse(open(path).read())
    assert C.c_int.in_dll(lib, "vops_count").value == len(ops)
    table = (VopDes * len(ops)).in_dll(lib, "vops")
    for o, d in zip(ops, table):
        assert d.vopname.decode() == f"fused{o.number}"
        assert d.arg_num == len(o.params) and d.result_num == 1
        assert [d.arg_type[i] for i in range(d.arg_num)] == [TYPE[t] for t in o.params]
        assert d.result_type[0] == TYPE[o.result]
        assert d.cvl_desc == len(o.params)                 # ElwiseK
        if o.reduce:
            r = o.reduce[1]
            assert d.result_len[0] == 1 << r and d.arg_len[r] == 0x20 and d.arg_for_scratch[0] == r + 1
        else:
            assert d.result_len[0] == (0x20 if o.params[0] == "SEGDES" else 0x1)
        for k in ("", "_scratch", "_inplace"):
            getattr(lib, f"fused{o.number}{k}")
