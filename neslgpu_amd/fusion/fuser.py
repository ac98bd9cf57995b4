"""VCODE fusion: .vcode -> .fcode + super-op definitions.

The counterpart of the reference fusion compiler's front end
(nesl/fusion/main.sml:33-80: parse, inline, convert, reduce, analyze
sizes, fuse, super-op conversion, VCODE generation), written for this
package's code generator (codegen.py), which turns the super-op definitions
into HIP.  The passes:

  parse     tokens and instructions of the stack language
            (nesl/vcode/grammar.yy, tokens.flex; comments, strings and
            vector constants are single tokens)
  inline    calls to small non-recursive functions are replaced by the
            callee's body (a VCODE call shares the caller's stack, so the
            body stands in for CALL as it is) -- nesl/fusion/inliner.sml
  convert   each straight-line region (between calls, branches, timers and
            I/O) is executed symbolically: stack slots become value nodes,
            COPY / POP / CPOP only move references, and every instruction
            becomes a node over the nodes it reads -- convert.sml
  reduce    nodes that no instruction and no stack slot at the region's
            end uses are dropped, unless they have effects (RAND) --
            reduce.sml
  sizes     segment descriptors known to hold one segment (MAKE_SEGDES of
            a LENGTH or of a scalar) are marked -- analyze-sizes.sml
  fuse      an elementwise node absorbs the elementwise nodes and the
            DISTs of scalar constants it reads; a BPERMUTE whose result
            feeds an elementwise node is absorbed as a gather; a +_REDUCE,
            MAX_REDUCE or MIN_REDUCE absorbs the elementwise tree it
            reduces.  Like the reference (fuse.sml:93-150), a node read by
            several fused trees is recomputed in each of them, and nothing
            is fused across a timer (fuse.sml:176-179) -- the regions end
            there
  super-op  every fused tree becomes a definition over its leaves
            ($0, $1, ... in order of first appearance), identical trees
            share one definition -- super-op.sml
  codegen   the region is written back as stack code: arguments are
            brought to the top with COPY (read again later) or CPOP (last
            read), each instruction or FUSED n follows, and the region ends
            with the stack laid out as the original left it -- vcodegen.sml

    python -m neslgpu_amd.fusion.fuser prog.vcode -o prog.fcode -s prog.superops
"""
from __future__ import annotations

import argparse
import re
import sys
from dataclasses import dataclass, field

# ------------------------------------------------------------------ opcodes
# Type tokens that follow each opcode (None: none) and the stack signature:
# argument types and result types, "G" standing for the instruction's type.
TYPES = ("INT", "FLOAT", "BOOL", "SEGDES", "CHAR")


@dataclass(frozen=True)
class OpSig:
    typed: bool
    args: tuple
    res: tuple
    kind: str            # "ew" elementwise, "vec" other vector op, "rand", "stack", "ctl", "io"


def _ew(args, res, typed=True):
    return OpSig(typed, args, res, "ew")


OPS = {
    "+": _ew("GG", "G"), "-": _ew("GG", "G"), "*": _ew("GG", "G"), "/": _ew("GG", "G"),
    "%": _ew("II", "I", False), "<": _ew("GG", "B"), "<=": _ew("GG", "B"), ">": _ew("GG", "B"),
    ">=": _ew("GG", "B"), "=": _ew("GG", "B"), "!=": _ew("GG", "B"),
    "LSHIFT": _ew("II", "I", False), "RSHIFT": _ew("II", "I", False),
    "NOT": _ew("G", "G"), "AND": _ew("GG", "G"), "OR": _ew("GG", "G"), "XOR": _ew("GG", "G"),
    "SELECT": _ew("BGG", "G"),
    "FLOOR": _ew("F", "I", False), "CEIL": _ew("F", "I", False), "TRUNC": _ew("F", "I", False),
    "ROUND": _ew("F", "I", False), "LOG": _ew("F", "F", False), "SQRT": _ew("F", "F", False),
    "EXP": _ew("F", "F", False), "SIN": _ew("F", "F", False), "COS": _ew("F", "F", False),
    "TAN": _ew("F", "F", False), "ASIN": _ew("F", "F", False), "ACOS": _ew("F", "F", False),
    "ATAN": _ew("F", "F", False), "SINH": _ew("F", "F", False), "COSH": _ew("F", "F", False),
    "TANH": _ew("F", "F", False), "I_TO_F": _ew("I", "F", False), "I_TO_B": _ew("I", "B", False),
    "B_TO_I": _ew("B", "I", False),
    "RAND": OpSig(False, "I", "I", "rand"),
    "PERMUTE": OpSig(True, "GIS", "G", "vec"), "DPERMUTE": OpSig(True, "GIGSS", "G", "vec"),
    "FPERMUTE": OpSig(True, "GIBSS", "G", "vec"), "BPERMUTE": OpSig(True, "GISS", "G", "vec"),
    "BFPERMUTE": OpSig(True, "GIBSS", "G", "vec"), "DFPERMUTE": OpSig(True, "GIBGSS", "G", "vec"),
    "EXTRACT": OpSig(True, "GIS", "G", "vec"), "REPLACE": OpSig(True, "GIGS", "G", "vec"),
    "DIST": OpSig(True, "GS", "G", "vec"), "INDEX": OpSig(False, "IIS", "I", "vec"),
    "RANK_UP": OpSig(True, "GS", "I", "vec"), "RANK_DOWN": OpSig(True, "GS", "I", "vec"),
    "PACK": OpSig(True, "GBS", "GS", "vec"), "LENGTH": OpSig(True, "G", "I", "vec"),
    "MAKE_SEGDES": OpSig(False, "I", "S", "vec"), "LENGTHS": OpSig(False, "S", "I", "vec"),
}
for _n in ("+_SCAN", "*_SCAN", "MAX_SCAN", "MIN_SCAN", "AND_SCAN", "OR_SCAN", "XOR_SCAN",
           "+_REDUCE", "*_REDUCE", "MAX_REDUCE", "MIN_REDUCE", "AND_REDUCE", "OR_REDUCE", "XOR_REDUCE"):
    OPS[_n] = OpSig(True, "GS", "G", "vec")
FIXED = {"I": "INT", "F": "FLOAT", "B": "BOOL", "S": "SEGDES"}
# instructions that end a straight-line region, with the number of type or
# operand tokens that follow them
BOUNDARY = {"CALL": 1, "IF": 0, "ELSE": 0, "ENDIF": 0, "RET": 0, "EXIT": 0, "START_TIMER": 0,
            "STOP_TIMER": 0, "READ": 1, "WRITE": 1, "FOPEN": 0, "FCLOSE": 0, "FREAD": 1, "FREAD_CHAR": 0,
            "FWRITE": 1, "SPAWN": 0, "SRAND": 1, "PAIR": 0, "UNPAIR": 0, "FUSED": 1}
STACK = ("COPY", "POP", "CPOP")
FUSE_REDUCES = ("+_REDUCE", "MAX_REDUCE", "MIN_REDUCE")


class FuseError(ValueError):
    pass


# ------------------------------------------------------------------ parse
_TOKEN = re.compile(r'\s*(\{[^}]*\}|"(?:[^"\\]|\\.)*"|\([^)]*\)|[^\s(){}"]+)', re.S)


def tokenize(text: str) -> list:
    out, pos = [], 0
    while True:
        m = _TOKEN.match(text, pos)
        if not m:
            if text[pos:].strip():
                raise FuseError(f"cannot read VCODE at {text[pos:pos + 40]!r}")
            return out
        pos = m.end()
        t = m.group(1)
        if not t.startswith("{"):
            out.append(t)


@dataclass
class Instr:
    op: str
    type: str = None
    a: object = None       # COPY/POP/CPOP count, CALL/FUNC label, CONST value, FUSED number
    b: object = None       # COPY/POP/CPOP depth

    def text(self) -> str:
        if self.op in STACK:
            return f"{self.op} {self.a} {self.b}"
        if self.op == "CONST":
            return f"CONST {self.type} {self.a}"
        if self.op in ("CALL", "FUNC", "FUSED"):
            return f"{self.op} {self.a}"
        return f"{self.op} {self.type}" if self.type else self.op


@dataclass
class Func:
    name: str
    body: list = field(default_factory=list)      # instructions, the final RET included


def parse(text: str) -> list:
    """[Func] of a VCODE program, in source order."""
    toks = tokenize(text)
    funcs, i = [], 0
    cur = None

    def need(k):
        if i + k >= len(toks):
            raise FuseError(f"truncated instruction {toks[i]}")
    while i < len(toks):
        t = toks[i]
        if t == "FUNC":
            need(1)
            cur = Func(toks[i + 1])
            funcs.append(cur)
            i += 2
            continue
        if cur is None:
            raise FuseError(f"instruction {t} outside a function")
        if t in STACK:
            need(2)
            cur.body.append(Instr(t, None, int(toks[i + 1]), int(toks[i + 2])))
            i += 3
        elif t == "CONST":
            need(2)
            cur.body.append(Instr(t, toks[i + 1], toks[i + 2]))
            i += 3
        elif t in ("CALL", "FUSED"):
            need(1)
            cur.body.append(Instr(t, None, toks[i + 1] if t == "CALL" else int(toks[i + 1])))
            i += 2
        elif t in BOUNDARY:
            k = BOUNDARY[t]
            cur.body.append(Instr(t, toks[i + 1] if k else None))
            i += 1 + k
        elif t in OPS:
            if OPS[t].typed:
                need(1)
                if toks[i + 1] not in TYPES:
                    raise FuseError(f"{t} needs a type, got {toks[i + 1]}")
                cur.body.append(Instr(t, toks[i + 1]))
                i += 2
            else:
                cur.body.append(Instr(t))
                i += 1
        else:
            raise FuseError(f"unknown instruction {t}")
    return funcs


def emit(funcs: list) -> str:
    out = []
    for f in funcs:
        out.append(f"FUNC {f.name}")
        out += [x.text() for x in f.body]
        out.append("")
    return "\n".join(out)


# ------------------------------------------------------------------ inline
def _calls(f: Func):
    return [x.a for x in f.body if x.op == "CALL"]


def inline(funcs: list, max_body: int = 12, max_rounds: int = 4) -> list:
    """Replace calls to small functions that call nothing themselves (so no
    recursion) by their bodies; functions nothing calls any more (the
    entry, the first function, aside) are dropped."""
    by = {f.name: f for f in funcs}
    for _ in range(max_rounds):
        small = {f.name for f in funcs[1:] if not _calls(f) and len(f.body) - 1 <= max_body
                 and all(x.op != "EXIT" for x in f.body) and f.body and f.body[-1].op == "RET"}
        changed = False
        for f in funcs:
            body = []
            for x in f.body:
                if x.op == "CALL" and x.a in small and x.a != f.name:
                    body += [Instr(y.op, y.type, y.a, y.b) for y in by[x.a].body[:-1]]
                    changed = True
                else:
                    body.append(x)
            f.body = body
        if not changed:
            break
    used = {funcs[0].name}
    work = [funcs[0].name]
    while work:
        for c in _calls(by[work.pop()]):
            if c not in used:
                used.add(c)
                work.append(c)
    return [f for f in funcs if f.name in used]


# ------------------------------------------------------------------ convert
class Node:
    """A value of a region: an instruction applied to nodes, or a slot of
    the stack the region started with (op None, `slot` its depth)."""
    __slots__ = ("op", "type", "itype", "args", "const", "slot", "out", "nres", "uses", "id", "fused", "one_seg",
                 "order")

    def __init__(self, op, type_, args=(), const=None, slot=None, out=0, nres=1, itype=None):
        # type: of the value; itype: the instruction's type token
        self.op, self.type, self.itype, self.args, self.const, self.slot = op, type_, itype, list(args), const, slot
        self.out, self.nres = out, nres          # which result of a multi-result instruction
        self.uses, self.id, self.fused, self.one_seg, self.order = 0, -1, None, False, 0


@dataclass
class Region:
    instrs: list
    slots: list = field(default_factory=list)    # live-in nodes, deepest first
    final: list = field(default_factory=list)    # the stack at the end (below: untouched)
    ops: list = field(default_factory=list)      # nodes in instruction order


def convert(instrs: list) -> Region:
    reg = Region(instrs)
    stack = []                # visible part of the stack, deepest first

    def reach(depth):         # make the slot `depth` below the top visible
        while len(stack) <= depth:
            n = Node(None, None, slot=len(reg.slots))
            reg.slots.append(n)
            stack.insert(0, n)

    for order, x in enumerate(instrs):
        if x.op in STACK:
            cnt, dep = x.a, x.b
            reach(cnt + dep - 1)
            lo = len(stack) - dep - cnt
            seg = stack[lo:lo + cnt]
            if x.op == "COPY":
                stack.extend(seg)
            elif x.op == "POP":
                del stack[lo:lo + cnt]
            else:
                del stack[lo:lo + cnt]
                stack.extend(seg)
            continue
        if x.op == "CONST":
            n = Node("CONST", x.type, const=x.a)
            n.order = order
            reg.ops.append(n)
            stack.append(n)
            continue
        sig = OPS[x.op]
        k = len(sig.args)
        reach(k - 1)
        args = stack[len(stack) - k:]
        del stack[len(stack) - k:]
        ty = x.type
        for a, want in zip(args, sig.args):
            t = ty if want == "G" else FIXED[want]
            if a.type is None and a.op is None:
                a.type = t                       # a live-in slot takes the type it is read as
        res = []
        for j, r in enumerate(sig.res):
            n = Node(x.op, ty if r == "G" else FIXED[r], args, out=j, nres=len(sig.res), itype=ty)
            n.order = order
            res.append(n)
        if len(res) > 1:
            for n in res:
                n.const = res                     # siblings of a multi-result instruction
        reg.ops.extend(res)
        stack.extend(res)
    reg.final = stack
    return reg


# ------------------------------------------------------------------ sizes, fuse
def _is_ew(n):
    return n.op is not None and n.op in OPS and OPS[n.op].kind == "ew"


def _is_scalar_const(n):
    return n.op == "CONST" and n.type in ("INT", "FLOAT", "BOOL") and not str(n.const).startswith("(")


def _is_flat_dist(n):
    return n.op == "DIST" and _is_scalar_const(n.args[0])


def analyze_sizes(reg: Region):
    for n in reg.ops:
        if n.op == "MAKE_SEGDES":
            src = n.args[0]
            n.one_seg = src.op == "LENGTH" or _is_scalar_const(src)


def fuse(reg: Region, reductions: bool = True, gathers: bool = True):
    """Mark every node that roots a fused tree (n.fused = True); the trees
    are the nodes' elementwise ancestry as absorbed() describes."""
    for n in reg.ops:
        n.fused = None
    for n in reg.ops:
        if _is_ew(n) and any(absorbable(a, gathers) for a in n.args):
            n.fused = True
        elif _is_flat_dist(n):
            n.fused = True
        elif reductions and n.op in FUSE_REDUCES and n.type in ("INT", "FLOAT") and \
                (_is_ew(n.args[0]) or _is_flat_dist(n.args[0]) or (gathers and n.args[0].op == "BPERMUTE")):
            n.fused = True


def absorbable(a, gathers=True):
    return _is_ew(a) or _is_flat_dist(a) or (gathers and a.op == "BPERMUTE" and a.type in ("INT", "FLOAT", "BOOL"))


# ------------------------------------------------------------------ super-ops
class SuperOps:
    """Definitions of the fused trees, deduplicated by text."""

    def __init__(self):
        self.defs, self.index, self.uses = [], {}, []

    def add(self, params, body) -> int:
        key = (tuple(params), body)
        if key not in self.index:
            self.index[key] = len(self.defs)
            self.defs.append((params, body))
            self.uses.append(0)
        k = self.index[key]
        self.uses[k] += 1
        return k

    def text(self) -> str:
        lines = []
        for k, (params, body) in enumerate(self.defs):
            ps = ", ".join(f"${i} : {t}" for i, t in enumerate(params))
            lines.append(f"fused OP{k}#{self.uses[k]} ({ps}) = {body}")
        return "\n".join(lines) + ("\n" if lines else "")


def tree(root: Node, gathers=True, order=None):
    """(body text, leaves, types of the leaves).  Leaves are numbered in
    order of first appearance, or as `order` lists them.  With no `order`,
    the leaves are then put in the order the fused-library ABI needs (see
    _abi_order) and the tree is written again in that order."""
    leaves, types, roles = [], [], []

    def leaf(n, t, role="vec"):
        if order is not None:
            for i, m in enumerate(order):
                if m is n:
                    if not any(x is n for x in leaves):
                        leaves.append(n)
                        types.append(t)
                        roles.append(role)
                    return f"${i}"
            raise FuseError("a leaf missing from the given order")
        for i, m in enumerate(leaves):
            if m is n:
                return f"${i}"
        leaves.append(n)
        types.append(t)
        roles.append(role)
        return f"${len(leaves) - 1}"

    def go(n, want_t, top):
        if not top and not absorbable(n, gathers):
            return leaf(n, want_t or n.type)      # the type the reading instruction takes it as
        if n.op == "DIST":
            c, sd = n.args
            return f"(DIST {n.itype} @ (CONST {c.type} {c.const}) {leaf(sd, 'SEGDES', 'dst')})"
        if n.op == "BPERMUTE":
            src, idx, sds, sdd = n.args
            s_src = leaf(src, n.itype, "src")
            s_idx = go(idx, "INT", False)
            return (f"(GATHER {n.itype} @ {s_src} {s_idx} {leaf(sds, 'SEGDES', 'srcsd')} "
                    f"{leaf(sdd, 'SEGDES', 'dst')})")
        sig = OPS[n.op]
        parts = []
        for a, w in zip(n.args, sig.args):
            parts.append(go(a, n.itype if w == "G" else FIXED[w], False))
        head = f"{n.op} {n.itype}" if sig.typed else n.op
        return f"({head} @ {' '.join(parts)})"

    if root.op in FUSE_REDUCES:
        v, sd = root.args
        inner = go(v, root.itype, True)
        body = f"({root.op} {root.itype} @ {inner} {leaf(sd, 'SEGDES', 'red')})"
    else:
        body = go(root, root.type, True)
    if order is None:
        new = _abi_order(leaves, types, roles)
        if any(a is not b for a, b in zip(new, leaves)):
            return tree(root, gathers, new)
    else:
        types = [types[[id(x) for x in leaves].index(id(n))] for n in order]
        leaves = list(order)
    return body, leaves, types


def _abi_order(leaves, types, roles):
    """Parameter order for the fused-library ABI (include/vcode_fused.h): the
    interpreter sizes the result by parameter 0 (a vector of the result's
    length, or a descriptor of it) and, for a reduction, by the reduced
    descriptor, which must be among the first five parameters."""
    idx = list(range(len(leaves)))
    first = [i for i in idx if roles[i] == "vec" and types[i] != "SEGDES"]
    if not first:
        first = [i for i in idx if roles[i] == "dst"]
    red = [i for i in idx if roles[i] == "red"]
    head = first[:1] + [i for i in red if i not in first[:1]]
    rest = [i for i in idx if i not in head]
    return [leaves[i] for i in head + rest]


# ------------------------------------------------------------------ codegen
def codegen(reg: Region, sops: SuperOps, gathers=True) -> list:
    """Stack code for the region: its live (not absorbed) nodes in order,
    fused roots as FUSED n, ending with the stack of the original region."""
    # what each emitted node reads
    reads, trees = {}, {}
    for n in reg.ops:
        if n.fused:
            trees[id(n)] = tree(n, gathers)
            reads[id(n)] = trees[id(n)][1]
        elif n.out == 0:
            reads[id(n)] = list(n.args)
    # liveness: keep a node when the final stack or a kept node reads it
    keep = set()
    for n in reg.final:
        keep.add(id(n))
    for n in reversed(reg.ops):
        if n.out != 0:
            continue
        sib = n.const if n.nres > 1 else [n]
        if any(id(s) in keep for s in sib) or n.op == "RAND":
            for s in sib:
                keep.add(id(s))
            for a in reads[id(n)]:
                keep.add(id(a))
    emitted = [n for n in reg.ops if n.out == 0 and id(n) in keep]
    for n in emitted:
        if n.fused:
            body, _, types = trees[id(n)]
            n.id = sops.add(types, body)
    # remaining reads of each node, for COPY (read again) against CPOP (last read)
    left = {}
    for n in emitted:
        for a in reads[id(n)]:
            left[id(a)] = left.get(id(a), 0) + 1
    for n in reg.final:
        left[id(n)] = left.get(id(n), 0) + 1
    out = []
    stack = list(reversed(reg.slots))    # deepest first, as the region found them

    floor = [0]                          # bottom slots that stay where they are

    def fetch(n, reserved):
        # bring n to the top, above the `reserved` values already fetched:
        # CPOP on its last read (unless it lies in the bottom part that
        # stays), else COPY
        for d in range(reserved, len(stack)):
            if stack[len(stack) - 1 - d] is n:
                break
        else:
            raise FuseError("a value the region reads is not on its stack")
        left[id(n)] -= 1
        if left[id(n)] <= 0 and len(stack) - 1 - d >= floor[0]:
            if d:
                out.append(Instr("CPOP", None, 1, d))
                stack.append(stack.pop(len(stack) - 1 - d))
        else:
            out.append(Instr("COPY", None, 1, d))
            stack.append(n)

    for n in emitted:
        for r, a in enumerate(reads[id(n)]):
            fetch(a, r)
        k = len(reads[id(n)])
        del stack[len(stack) - k:]
        if n.fused:
            out.append(Instr("FUSED", None, n.id))
            stack.append(n)
        elif n.op == "CONST":
            out.append(Instr("CONST", n.type, n.const))
            stack.append(n)
        else:
            out.append(Instr(n.op, n.itype if OPS[n.op].typed else None))
            if n.nres > 1:
                stack.extend(n.const)
            else:
                stack.append(n)
    # the final layout: the part already in place at the bottom stays; the
    # rest is built on top, then what lies between is popped
    keep_n = 0
    while keep_n < min(len(stack), len(reg.final)) and stack[keep_n] is reg.final[keep_n]:
        keep_n += 1
    for n in reg.final[:keep_n]:
        left[id(n)] -= 1
    floor[0] = keep_n
    for r, n in enumerate(reg.final[keep_n:]):
        fetch(n, r)
    top = len(reg.final) - keep_n
    below = len(stack) - len(reg.final)
    if below:
        out.append(Instr("POP", None, below, top))
    return out


def _pop_merge(code: list) -> list:
    """Peephole: `CPOP 1 0` is a no-op; `COPY 1 d` followed by `POP 1 0` too."""
    out = []
    for x in code:
        if x.op == "CPOP" and x.b == 0:
            continue
        if x.op == "POP" and x.a == 1 and x.b == 0 and out and out[-1].op == "COPY" and out[-1].a == 1:
            out.pop()
            continue
        out.append(x)
    return out


# ------------------------------------------------------------------ driver
def fuse_program(text: str, do_inline: bool = True, reductions: bool = True, gathers: bool = True):
    """(fused VCODE text, super-op definitions text)."""
    funcs = parse(text)
    if do_inline:
        funcs = inline(funcs)
    sops = SuperOps()
    for f in funcs:
        body, cur = [], []

        def flush():
            if cur:
                reg = convert(cur)
                analyze_sizes(reg)
                fuse(reg, reductions, gathers)
                body.extend(_pop_merge(codegen(reg, sops, gathers)))
                cur.clear()
        for x in f.body:
            if x.op in BOUNDARY:
                flush()
                body.append(x)
            else:
                cur.append(x)
        flush()
        f.body = body
    return emit(funcs), sops.text()


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("vcode")
    p.add_argument("-o", "--fcode", required=True)
    p.add_argument("-s", "--superops", required=True)
    p.add_argument("--no-inline", action="store_true")
    a = p.parse_args(argv)
    fcode, sops = fuse_program(open(a.vcode).read(), do_inline=not a.no_inline)
    with open(a.fcode, "w") as f:
        f.write(fcode)
    with open(a.superops, "w") as f:
        f.write(sops)


if __name__ == "__main__":
    sys.exit(main())
