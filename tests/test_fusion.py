"""Fusion back end without a GPU: super-op parsing and typing, generated
source, and the descriptor tables of the built fused libraries (read with
ctypes; nothing is launched)."""
import ctypes as C
import glob
import os

import pytest

from neslgpu_amd.build import FUSED_DIR, GEN_DIR, PROGRAMS_DIR, SUPEROPS_DIR
from neslgpu_amd.fusion import codegen, superops as S

DEFS = sorted(glob.glob(os.path.join(SUPEROPS_DIR, "*.superops")) + glob.glob(os.path.join(PROGRAMS_DIR, "*.superops")))
# the definitions the fuser wrote at build time, each beside its library
GEN_DEFS = sorted(glob.glob(os.path.join(GEN_DIR, "*.superops")))


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


def interpreter_accepts(o, arg_len, result_len):
    """The length rules of a definition, applied as the interpreter applies
    them (vcode/machine.cpp check_args, result_len) to arguments of
    consistent sizes: n elements; every descriptor over the elements
    describes n of them in a segment count of its own; a gathered source has
    a length of its own, which its descriptor describes.  The rules must
    accept them and give the result n elements (one per segment of the
    reduced descriptor for a reduction)."""
    n = 1000
    gsrc = S.gather_sources(o.body)
    gseg = {v: k for k, v in gsrc.items()}
    args = []
    for i, t in enumerate(o.params):
        if i in gsrc:
            args.append((5000 + i, None))                          # (len, seg_len)
        elif t == "SEGDES":
            args.append((7 + i, 5000 + gseg[i] if i in gseg else n))
        else:
            args.append((n, None))
    for k, r in enumerate(arg_len):
        for j in range(5):
            if r & (1 << j) and j < len(args) and args[k][0] != args[j][0]:
                return False
            if r & (0x20 << j) and j < len(args) and args[k][1] != args[j][0]:
                return False
    want = args[o.reduce[1]][0] if o.reduce else n
    for j in range(5):
        if result_len & (1 << j):
            return args[j][0] == want
    for j in range(5):
        if result_len & (0x20 << j):
            return args[j][1] == want
    return False


def test_length_rules_of_descriptor_only_definitions():
    """A reduction whose only argument is its descriptor (a +_REDUCE of a
    DIST of a constant) and an elementwise tree over two descriptors of the
    same elements in different segment counts."""
    ops = S.parse("fused OP0#1 ($0 : SEGDES) = (+_REDUCE INT @ (DIST INT @ (CONST INT 5) $0) $0)\n"
                  "fused OP1#1 ($0 : SEGDES, $1 : SEGDES) = "
                  "(+ INT @ (DIST INT @ (CONST INT 1) $0) (DIST INT @ (CONST INT 2) $1))\n"
                  "fused OP2#1 ($0 : INT, $1 : SEGDES, $2 : INT) = (+_REDUCE INT @ (* INT @ $0 $2) $1)\n")
    src = codegen.generate(ops, "x")
    rows = [l for l in src.splitlines() if l.strip().startswith('{VCODE_FUSED_TOKEN')]
    lens = [l.split("}, {")[1] for l in rows]
    assert lens == ["NONE", "NONE, NONE", "NONE, COMPAT1, AGREE1"]
    rule = {"NONE": 0, "COMPAT1": 0x20, "AGREE1": 0x1}
    for o, l, row in zip(ops, lens, rows):
        res = row.split("}, {")[3]
        res_rule = 1 << o.reduce[1] if o.reduce else 0x20
        assert res == (f"AGREE{o.reduce[1] + 1}" if o.reduce else "COMPAT1")
        assert interpreter_accepts(o, [rule[x] for x in l.split(", ")], res_rule)


@pytest.mark.parametrize("path", DEFS + GEN_DEFS, ids=lambda p: os.path.relpath(p, FUSED_DIR))
def test_built_library_tables(path):
    name = os.path.basename(path)[: -len(".superops")]
    so = os.path.join(GEN_DIR if path in GEN_DEFS else FUSED_DIR, name + ".so")
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
            assert d.result_len[0] == 1 << r and d.arg_for_scratch[0] == r + 1
        else:
            assert d.result_len[0] == (0x20 if o.params[0] == "SEGDES" else 0x1)
        arg_len = [d.arg_len[i] for i in range(d.arg_num)]
        assert interpreter_accepts(o, arg_len, d.result_len[0]), (o.number, arg_len)
        for k in ("", "_scratch", "_inplace"):
            getattr(lib, f"fused{o.number}{k}")
