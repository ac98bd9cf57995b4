"""Super-operator definitions: parser and type checker.

A fused VCODE program (.fcode) replaces runs of elementwise instructions
with `FUSED n`; super-op n is defined by one line of the form the reference
fusion compiler writes (its generator is nesl/fusion/super-op-gen.sml:42-480;
examples trail each bench/NESL/*.cu):

    fused OP3#2 ($0 : INT, $1 : SEGDES) = (/ INT @ $0 (DIST INT @ (CONST INT 2) $1))

The right-hand side is an expression tree over the arguments `$i`:
`(OP [TYPE] @ e1 e2 ...)` applies VCODE elementwise op OP at TYPE,
`(CONST TYPE v)` is a scalar, and `(DIST TYPE @ c $s)` spreads a scalar over
the elements segment descriptor $s describes.  Segment-descriptor arguments
only give the length.

Two forms go beyond the reference fusion compiler, which fuses elementwise
code only.  At the root, `(+_REDUCE TYPE @ e $s)` (also MAX_REDUCE,
MIN_REDUCE) reduces e over each segment of $s: one result per segment, the
products of a sparse matrix-vector product summed per row without writing
them.  And `(GATHER TYPE @ $src e $s)` reads $src at the index e evaluates to
(an unsegmented BPERMUTE folded into the fused op; $s is the one-segment
descriptor of $src, whose length bounds the index: an index outside
[0, length) gives 0).  $src need not have the length of the other arguments.
The four-argument form `(GATHER TYPE @ $src e $ss $sd)` is a whole segmented
BPERMUTE (ref nesl/cvl/serial/vprims.c:416-467): element k of the result,
in segment j of $sd, reads $src at e plus the start of segment j of $ss (an
index outside [0, length of segment j of $ss) gives 0).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

TYPES = ("INT", "FLOAT", "BOOL", "SEGDES")


@dataclass
class Arg:
    index: int


@dataclass
class Const:
    type: str
    value: str


@dataclass
class Apply:
    op: str
    type: str | None
    args: list = field(default_factory=list)


@dataclass
class SuperOp:
    number: int
    uses: int
    params: list            # types of $0, $1, ...
    body: object
    result: str = ""        # result type, filled by typecheck
    text: str = ""
    reduce: tuple = None    # (operator, segd parameter) when the root is a segmented reduction


# result type of each elementwise op, given its operand type
_SAME = {"+", "-", "*", "/", "%", "NOT", "AND", "OR", "XOR", "MAX", "MIN"}
_BOOL = {"<", "<=", ">", ">=", "=", "!="}
_FIXED = {
    "LSHIFT": "INT", "RSHIFT": "INT", "B_TO_I": "INT", "I_TO_F": "FLOAT", "I_TO_B": "BOOL",
    "ROUND": "INT", "FLOOR": "INT", "CEIL": "INT", "TRUNC": "INT",
    "LOG": "FLOAT", "SQRT": "FLOAT", "EXP": "FLOAT", "SIN": "FLOAT", "COS": "FLOAT", "TAN": "FLOAT",
    "ASIN": "FLOAT", "ACOS": "FLOAT", "ATAN": "FLOAT", "SINH": "FLOAT", "COSH": "FLOAT", "TANH": "FLOAT",
}
# operand type of the ops whose instruction carries no type
_OPERAND = {
    "LSHIFT": "INT", "RSHIFT": "INT", "B_TO_I": "BOOL", "I_TO_F": "INT", "I_TO_B": "INT", "%": "INT",
    "ROUND": "FLOAT", "FLOOR": "FLOAT", "CEIL": "FLOAT", "TRUNC": "FLOAT",
    "LOG": "FLOAT", "SQRT": "FLOAT", "EXP": "FLOAT", "SIN": "FLOAT", "COS": "FLOAT", "TAN": "FLOAT",
    "ASIN": "FLOAT", "ACOS": "FLOAT", "ATAN": "FLOAT", "SINH": "FLOAT", "COSH": "FLOAT", "TANH": "FLOAT",
}

REDUCES = ("+_REDUCE", "MAX_REDUCE", "MIN_REDUCE")

_HEAD = re.compile(r"^fused\s+OP(\d+)#(\d+)\s*\((.*?)\)\s*=\s*(.*)$")


class DefinitionError(ValueError):
    pass


def _tokens(s: str):
    return re.findall(r"\(|\)|@|[^\s()@]+", s)


def _parse_expr(toks, pos):
    t = toks[pos]
    if t.startswith("$"):
        return Arg(int(t[1:])), pos + 1
    if t != "(":
        raise DefinitionError(f"unexpected token {t!r}")
    op = toks[pos + 1]
    pos += 2
    if op == "CONST":
        typ, val, close = toks[pos], toks[pos + 1], toks[pos + 2]
        if close != ")" or typ not in TYPES:
            raise DefinitionError("malformed CONST")
        return Const(typ, val), pos + 3
    typ = None
    if toks[pos] in TYPES:
        typ = toks[pos]
        pos += 1
    if toks[pos] != "@":
        raise DefinitionError(f"expected @ after {op}")
    pos += 1
    args = []
    while toks[pos] != ")":
        e, pos = _parse_expr(toks, pos)
        args.append(e)
    return Apply(op, typ, args), pos + 1


def parse_line(line: str) -> SuperOp:
    m = _HEAD.match(line.strip())
    if not m:
        raise DefinitionError(f"not a super-op definition: {line!r}")
    params = []
    if m.group(3).strip():
        for i, p in enumerate(m.group(3).split(",")):
            name, typ = [x.strip() for x in p.split(":")]
            if name != f"${i}" or typ not in TYPES:
                raise DefinitionError(f"bad parameter {p!r}")
            params.append(typ)
    toks = _tokens(m.group(4))
    body, pos = _parse_expr(toks, 0)
    if pos != len(toks):
        raise DefinitionError("trailing tokens")
    op = SuperOp(int(m.group(1)), int(m.group(2)), params, body, text=line.strip())
    if isinstance(body, Apply) and body.op in REDUCES:
        if len(body.args) != 2 or not isinstance(body.args[1], Arg) or params[body.args[1].index] != "SEGDES":
            raise DefinitionError(f"{body.op} needs an expression and a segment descriptor")
        op.reduce = (body.op, body.args[1].index)
        op.body = body.args[0]
        t = typecheck(op)
        if body.type and body.type != t:
            raise DefinitionError(f"{body.op} {body.type} of a {t} expression")
        if t not in ("INT", "FLOAT"):
            raise DefinitionError(f"{body.op} of {t} is not supported")
        op.result = t
        return op
    op.result = typecheck(op)
    return op


def typecheck(op: SuperOp) -> str:
    def ty(e):
        if isinstance(e, Arg):
            if e.index >= len(op.params):
                raise DefinitionError(f"${e.index} is not a parameter")
            t = op.params[e.index]
            if t == "SEGDES":
                raise DefinitionError("a segment descriptor used as a value")
            return t
        if isinstance(e, Const):
            return e.type
        if e.op == "DIST":
            if len(e.args) != 2 or not isinstance(e.args[1], Arg) or op.params[e.args[1].index] != "SEGDES":
                raise DefinitionError("DIST needs a scalar and a segment descriptor")
            if not isinstance(e.args[0], Const):
                raise DefinitionError("only constants can be distributed inside a fused op")
            return ty(e.args[0])
        if e.op == "GATHER":
            if len(e.args) not in (3, 4) or not isinstance(e.args[0], Arg) or \
                    not all(isinstance(x, Arg) for x in e.args[2:]):
                raise DefinitionError("GATHER needs $src, an index and one or two descriptors")
            if any(op.params[x.index] != "SEGDES" for x in e.args[2:]) or ty(e.args[1]) != "INT":
                raise DefinitionError("GATHER index must be INT and its descriptors segment descriptors")
            t = ty(e.args[0])
            if e.type and e.type != t:
                raise DefinitionError("GATHER type")
            return t
        if e.op == "SELECT":
            f, a, b = (ty(x) for x in e.args)
            if f != "BOOL" or a != b:
                raise DefinitionError("SELECT types")
            return a
        if e.op == "RAND":
            raise DefinitionError("RAND cannot be fused")
        if e.op in REDUCES:
            raise DefinitionError(f"{e.op} can only be the root of a fused op")
        ts = [ty(x) for x in e.args]
        want = e.type or _OPERAND.get(e.op)
        if want is None:
            raise DefinitionError(f"{e.op} needs a type")
        if any(t != want for t in ts):
            raise DefinitionError(f"{e.op} {want} applied to {ts}")
        if e.op in _SAME:
            return want
        if e.op in _BOOL:
            return "BOOL"
        if e.op in _FIXED:
            return _FIXED[e.op]
        raise DefinitionError(f"unknown elementwise op {e.op}")
    return ty(op.body)


def gather_sources(e, acc=None):
    """{src index: segd index} of the GATHER forms of an expression (the
    descriptor of the source)."""
    acc = {} if acc is None else acc
    if isinstance(e, Apply):
        if e.op == "GATHER":
            acc[e.args[0].index] = e.args[2].index
        for x in e.args:
            gather_sources(x, acc)
    return acc


def gather_dests(e, acc=None):
    """Indices of the destination descriptors of four-argument GATHERs."""
    acc = set() if acc is None else acc
    if isinstance(e, Apply):
        if e.op == "GATHER" and len(e.args) == 4:
            acc.add(e.args[3].index)
        for x in e.args:
            gather_dests(x, acc)
    return acc


def parse(text: str) -> list:
    """All definitions of a text (one may wrap over several lines), ordered
    by number; numbers must be 0..n-1."""
    defs = []
    for l in text.splitlines():
        if l.strip().startswith("fused "):
            defs.append(l.strip())
        elif defs and l.strip():
            defs[-1] += " " + l.strip()
    ops = [parse_line(d) for d in defs]
    ops.sort(key=lambda o: o.number)
    if [o.number for o in ops] != list(range(len(ops))):
        raise DefinitionError("super-ops must be numbered 0..n-1 without gaps")
    return ops
