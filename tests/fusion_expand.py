"""Test helper: a fused program (.fcode + super-op definitions) written back
as plain VCODE, every `FUSED n` replaced by the stack code that evaluates
super-op n's tree one VCODE instruction at a time.  The expanded program
runs on any CVL library -- the reference serial CVL on the host included --
so what a fused program computes can be compared with what its unfused
source prints, without the GPU libraries generated from the definitions.

Arguments of `FUSED n` are on the stack with $0 deepest; the expansion
pushes the tree's value on top of them and then pops them from under it.
"""
from neslgpu_amd.fusion import superops as S

_UNTYPED = {"%", "LSHIFT", "RSHIFT", "B_TO_I", "I_TO_F", "I_TO_B", "ROUND", "FLOOR", "CEIL", "TRUNC",
            "LOG", "SQRT", "EXP", "SIN", "COS", "TAN", "ASIN", "ACOS", "ATAN", "SINH", "COSH", "TANH"}


def expand_op(op: S.SuperOp) -> list:
    k = len(op.params)
    code = []
    pushed = [0]

    def arg(i):
        code.append(f"COPY 1 {k - 1 - i + pushed[0]}")
        pushed[0] += 1

    def go(e):
        if isinstance(e, S.Arg):
            arg(e.index)
            return
        if isinstance(e, S.Const):
            code.append(f"CONST {e.type} {e.value}")
            pushed[0] += 1
            return
        if e.op == "DIST":
            go(e.args[0])
            arg(e.args[1].index)
            code.append(f"DIST {e.type}")
            pushed[0] -= 1
            return
        if e.op == "GATHER":
            src, idx = e.args[0], e.args[1]
            arg(src.index)
            go(idx)
            if len(e.args) != 4:
                raise ValueError("expansion of the one-descriptor GATHER form")
            arg(e.args[2].index)
            arg(e.args[3].index)
            code.append(f"BPERMUTE {e.type or op.params[src.index]}")
            pushed[0] -= 3
            return
        for x in e.args:
            go(x)
        code.append(e.op if e.op in _UNTYPED else f"{e.op} {e.type}")
        pushed[0] -= len(e.args) - 1

    if op.reduce:
        rop, r = op.reduce
        go(op.body)
        arg(r)
        code.append(f"{rop} {op.result}")
        pushed[0] -= 1
    else:
        go(op.body)
    assert pushed[0] == 1, op.text
    if k:
        code.append(f"POP {k} 1")
    return code


def expand_program(fcode: str, superops_text: str) -> str:
    ops = {o.number: o for o in S.parse(superops_text)}
    out = []
    for line in fcode.splitlines():
        s = line.split()
        if len(s) == 2 and s[0] == "FUSED":
            out += expand_op(ops[int(s[1])])
        else:
            out.append(line)
    return "\n".join(out) + "\n"
