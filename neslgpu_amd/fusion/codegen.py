"""HIP code generator for fused VCODE operations.

From super-op definitions (superops.py) it writes one HIP source for gfx950
and builds it into a shared library with the fused-library ABI of
include/vcode_fused.h, the counterpart of the CUDA the reference fusion
compiler emits (nesl/fusion/super-op-gen.sml:42-480).  Per super-op:

  - one grid-stride kernel that loads each vector argument once, evaluates
    the expression tree in registers with the CVL's own elementwise functors
    (csrc/elwise_ops.hpp, so a fused op rounds exactly like the CVL calls it
    replaces; FLOAT is double throughout) and stores the result once:
    (k + 1) * width bytes of HBM traffic per element instead of the
    unfused sequence's reads and writes of every intermediate vector;
  - the host entry  fusedN(vec_p d, vec_p s0.., int len, vec_p scratch)  on
    the CVL library's stream (cvlhip_get_stream), fusedN_scratch and
    fusedN_inplace (argument 0 may be overwritten when it is a vector);
  - its vopdes_t descriptor in `vops` and its triple in `cvl_funs`.

    python -m neslgpu_amd.fusion defs.superops -o libfused.so [--hip out.hip]
"""
from __future__ import annotations

import os
import subprocess
import sys

from . import superops as S

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CVLHIP_ARCH", "gfx950")

CTYPE = {"INT": "int", "BOOL": "int", "FLOAT": "double"}
VTYPE = {"INT": "Int", "BOOL": "Bool", "FLOAT": "Float", "SEGDES": "Segdes"}

# (op, operand type) -> functor of csrc/elwise_ops.hpp
_F = {
    ("+", "INT"): "AddZf", ("+", "FLOAT"): "AddDf", ("-", "INT"): "SubZf", ("-", "FLOAT"): "SubDf",
    ("*", "INT"): "MulZf", ("*", "FLOAT"): "MulDf", ("/", "INT"): "DivZf", ("/", "FLOAT"): "DivDf",
    ("%", "INT"): "ModZf", ("MAX", "INT"): "MaxZf", ("MAX", "FLOAT"): "MaxDf", ("MIN", "INT"): "MinZf",
    ("MIN", "FLOAT"): "MinDf",
    ("<", "INT"): "LesZf", ("<", "FLOAT"): "LesDf", ("<=", "INT"): "LeqZf", ("<=", "FLOAT"): "LeqDf",
    (">", "INT"): "GrtZf", (">", "FLOAT"): "GrtDf", (">=", "INT"): "GeqZf", (">=", "FLOAT"): "GeqDf",
    ("=", "INT"): "EqlZf", ("=", "FLOAT"): "EqlDf", ("=", "BOOL"): "EqlBf",
    ("!=", "INT"): "NeqZf", ("!=", "FLOAT"): "NeqDf", ("!=", "BOOL"): "NeqBf",
    ("LSHIFT", "INT"): "LshZf", ("RSHIFT", "INT"): "RshZf",
    ("NOT", "BOOL"): "NotBf", ("NOT", "INT"): "NotZf", ("AND", "BOOL"): "AndBf", ("AND", "INT"): "AndZf",
    ("OR", "BOOL"): "IorBf", ("OR", "INT"): "IorZf", ("XOR", "BOOL"): "XorBf", ("XOR", "INT"): "XorZf",
    ("B_TO_I", "BOOL"): "IntBf", ("I_TO_F", "INT"): "DblZf", ("I_TO_B", "INT"): "BooZf",
    ("ROUND", "FLOAT"): "RouDf", ("FLOOR", "FLOAT"): "FlrDf", ("CEIL", "FLOAT"): "CeiDf",
    ("TRUNC", "FLOAT"): "TrnDf", ("LOG", "FLOAT"): "LogDf", ("SQRT", "FLOAT"): "SqtDf",
    ("EXP", "FLOAT"): "ExpDf", ("SIN", "FLOAT"): "SinDf", ("COS", "FLOAT"): "CosDf",
    ("TAN", "FLOAT"): "TanDf", ("ASIN", "FLOAT"): "AsnDf", ("ACOS", "FLOAT"): "AcsDf",
    ("ATAN", "FLOAT"): "AtnDf", ("SINH", "FLOAT"): "SnhDf", ("COSH", "FLOAT"): "CshDf",
    ("TANH", "FLOAT"): "TnhDf",
}


def _literal(c: S.Const) -> str:
    if c.type == "FLOAT":
        return f"{float(c.value)!r}" if "e" not in repr(float(c.value)) else f"({float(c.value)!r})"
    if c.type == "BOOL":
        return "1" if c.value == "T" else "0"
    return f"({int(c.value)})"


def _expr(op: S.SuperOp, e) -> str:
    if isinstance(e, S.Arg):
        return f"a{e.index}"
    if isinstance(e, S.Const):
        return _literal(e)
    if e.op == "DIST":
        return _expr(op, e.args[0])
    if e.op == "GATHER":
        src, seg = e.args[0].index, e.args[2].index
        if len(e.args) == 4:                  # segmented: the segment of element ix in the destination
            return f"gather_seg(s{src}, v{seg}, v{e.args[3].index}, {_expr(op, e.args[1])}, ix)"
        return f"gather(s{src}, n{seg}, {_expr(op, e.args[1])})"
    if e.op == "SELECT":
        f, a, b = (_expr(op, x) for x in e.args)
        return f"({f} ? {a} : {b})"
    operand = e.type or S._OPERAND[e.op]
    fn = _F.get((e.op, operand))
    if fn is None:
        raise S.DefinitionError(f"no elementwise operator {e.op} {operand}")
    return f"cvlhip::{fn}()({', '.join(_expr(op, x) for x in e.args)})"


def _used_args(e, acc):
    """Arguments read elementwise (GATHER sources and segment descriptors
    are passed separately)."""
    if isinstance(e, S.Arg):
        acc.add(e.index)
    elif isinstance(e, S.Apply):
        if e.op == "DIST":
            return _used_args(e.args[0], acc)
        if e.op == "GATHER":
            return _used_args(e.args[1], acc)
        for x in e.args:
            _used_args(x, acc)
    return acc


_REDUCE_OP = {("+_REDUCE", "INT"): "AddZ", ("+_REDUCE", "FLOAT"): "AddD", ("MAX_REDUCE", "INT"): "MaxZ",
              ("MAX_REDUCE", "FLOAT"): "MaxD", ("MIN_REDUCE", "INT"): "MinZ", ("MIN_REDUCE", "FLOAT"): "MinD"}


def _reduce_op(o) -> list:
    """A super-op whose root reduces an expression over the segments of a
    descriptor: the CVL's own segmented reduction (segsched.hpp,
    seg_reduce_run: the schedule the descriptor names, the same association
    as *_re*) with an input loader that evaluates the expression, so the
    per-element values are never stored."""
    k, (rop, r) = o.number, o.reduce
    op = _REDUCE_OP[(rop, o.result)]
    rt = CTYPE[o.result]
    used = sorted(_used_args(o.body, set()))
    gsrc = S.gather_sources(o.body)
    gseg = sorted(set(gsrc.values()) | S.gather_dests(o.body))
    fields = [f"const {CTYPE[o.params[i]]} *s{i};" for i in used]
    fields += [f"const {CTYPE[o.params[i]]} *s{i};" for i in sorted(gsrc) if i not in used]
    fields += [f"const void *sd{j};" for j in gseg]
    L = [f"struct InFused{k} {{"]
    L += [f"    {f}" for f in fields]
    L += ["    template <class Op, int V>",
          f"    __device__ __forceinline__ void load(int64_t base, int64_t n, bool aligned, {rt} (&x)[V]) const {{"]
    for j in gseg:
        L.append(f"        const long long n{j} = *(const long long *)sd{j};")
        L.append(f"        const cvlhip::SegdView v{j} = cvlhip::SegdView::from_header(sd{j});")
        L.append(f"        (void)n{j};")
        L.append(f"        (void)v{j};")
    # Whole rows of aligned arguments: one vector load of V elements per
    # streamed argument (`aligned` is set by the host when every streamed
    # argument starts at a multiple of its vector size); gathered sources
    # stay per element.
    L += ["        if (aligned && base + V <= n) {"]
    for i in used:
        L.append(f"            const typename FusedVec<{CTYPE[o.params[i]]}, V>::type w{i} = "
                 f"*(const typename FusedVec<{CTYPE[o.params[i]]}, V>::type *)(s{i} + base);")
    L += ["#pragma unroll", "            for (int k = 0; k < V; k++) {",
          "                const long long ix = base + k;", "                (void)ix;"]
    for i in used:
        L.append(f"                const {CTYPE[o.params[i]]} a{i} = w{i}[k];")
    L += [f"                x[k] = ({rt})({_expr(o, o.body)});", "            }", "            return;", "        }"]
    L += ["#pragma unroll", "        for (int k = 0; k < V; k++) {", "            const int64_t i = base + k;",
          "            const long long ix = i;", "            (void)ix;", "            if (i < n) {"]
    for i in used:
        L.append(f"                const {CTYPE[o.params[i]]} a{i} = s{i}[i];")
    L += [f"                x[k] = ({rt})({_expr(o, o.body)});",
          "            } else {", "                x[k] = Op::neutral();", "            }", "        }", "    }", "};"]
    hparams = ["vec_p d"] + [f"vec_p s{i}" for i in range(len(o.params))] + ["int len", "vec_p scratch"]
    init = [f"(const {CTYPE[o.params[i]]} *)s{i}" for i in used]
    init += [f"(const {CTYPE[o.params[i]]} *)s{i}" for i in sorted(gsrc) if i not in used]
    init += [f"(const void *)s{j}" for j in gseg]
    L += [f'extern "C" void fused{k}({", ".join(hparams)}) {{',
          f"    using Op = cvlhip::{op};",
          "    constexpr int V = cvlhip::lane_values<Op>::value;",
          "    hipStream_t st = (hipStream_t)cvlhip_get_stream();"]
    for i in range(len(o.params)):
        if i not in used and i not in gsrc and i not in gseg and i != r:
            L.append(f"    (void)s{i};")
    L += [f"    const void *sd = (const void *)s{r};",
          f"    cvlhip::k_empty_identity_hdr<{rt}><<<256, 256, 0, st>>>(({rt} *)d, sd, Op::cvl_id());",
          f'    fused_check("fused{k}");',
          "    if (len <= 0) return;",
          "    const int aligned = " + (" && ".join(
              f"((unsigned long long)s{i} % (sizeof({CTYPE[o.params[i]]}) * V)) == 0" for i in used) or "0") + ";",
          "    // the CVL's own segmented reduction, under the schedule the descriptor",
          "    // names (the same tiles and bits as the CVL call on the same descriptor)",
          f"    cvlhip::seg_reduce_run<Op, InFused{k}>(InFused{k}{{{', '.join(init)}}}, ({rt} *)d, (int64_t)len,",
          "                                          cvlhip::SegdView(), sd, true, scratch, aligned != 0,",
          "                                          (unsigned *)cvlhip_device_error_word(), -1, st);",
          f'    fused_check("fused{k}");',
          "}",
          f'extern "C" int fused{k}_scratch(int n, int m) {{',
          "    (void)m;",
          f"    return (int)((cvlhip::sched_scratch_bytes<cvlhip::{op}, true>(n) + 7) / 8);",
          "}",
          f'extern "C" unsigned fused{k}_inplace(void) {{ return 0u; }}', ""]
    return L


def generate(ops, library_name: str = "fused") -> str:
    """HIP source of a fused-operation library."""
    out = [f"// {library_name}: fused VCODE operations generated by neslgpu_amd.fusion for gfx950.",
           "// Each kernel evaluates one super-op per element; definitions:"]
    out += [f"//   {o.text}" for o in ops]
    out += ['#include <hip/hip_runtime.h>', '#include <stdio.h>', '#include <stdlib.h>',
            '#include "cvl.h"', '#include "cvl_hip_ext.h"', '#include "vcode_fused.h"',
            '#include "elwise_ops.hpp"', '#include "segsched.hpp"', "",
            "static int fused_grid(long long n) {",
            "    long long g = (n + 1023) / 1024;",
            "    return g > 16384 ? 16384 : (g < 1 ? 1 : (int)g);",
            "}", "",
            "// V consecutive elements of one argument, read with one vector load.",
            "template <class T, int V> struct FusedVec { typedef T type __attribute__((ext_vector_type(V))); };", "",
            "template <class T>",
            "__device__ __forceinline__ T gather(const T *src, long long n, int i) {",
            "    return i >= 0 && (long long)i < n ? src[i] : (T)0;",
            "}", "",
            "// Segmented back-permute of element k: its segment j in the destination",
            "// descriptor vd, index i within segment j of the source descriptor vs.",
            "template <class T>",
            "__device__ __forceinline__ T gather_seg(const T *src, const cvlhip::SegdView &vs,",
            "                                        const cvlhip::SegdView &vd, int i, long long k) {",
            "    if (vd.m == 1) return i >= 0 && (long long)i < vs.n ? src[i] : (T)0;",
            "    const int j = vd.seg_of(k);",
            "    return i >= 0 && i < vs.lens[j] ? src[vs.offs[j] + i] : (T)0;",
            "}", "",
            "static void fused_check(const char *what) {",
            "    hipError_t e = hipGetLastError();",
            "    if (e != hipSuccess) {",
            '        fprintf(stderr, "%s: %s\\n", what, hipGetErrorString(e));',
            "        exit(1);",
            "    }",
            "}", ""]
    for o in ops:
        if o.reduce:
            out += _reduce_op(o)
            continue
        k = o.number
        used = sorted(_used_args(o.body, set()))
        gsrc = S.gather_sources(o.body)
        gdesc = sorted(set(gsrc.values()) | S.gather_dests(o.body))
        rt = CTYPE[o.result]
        params = [f"{rt} *d"]                  # may alias argument 0 (in place)
        params += [f"const {CTYPE[o.params[i]]} *s{i}" for i in used]
        params += [f"const {CTYPE[o.params[i]]} *s{i}" for i in sorted(gsrc) if i not in used]
        params += [f"const int *sd{j}" for j in gdesc]
        params.append("long long n")
        out.append(f"__global__ void __launch_bounds__(256) k_fused{k}({', '.join(params)}) {{")
        out.append("    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;")
        out.append("    const long long stride = (long long)gridDim.x * blockDim.x;")
        for j in gdesc:                           # shape from the descriptor's header
            out.append(f"    const long long n{j} = *(const long long *)sd{j};")
            out.append(f"    const cvlhip::SegdView v{j} = cvlhip::SegdView::from_header(sd{j});")
            out.append(f"    (void)n{j};")
            out.append(f"    (void)v{j};")
        out.append("    for (; i < n; i += stride) {")
        out.append("        const long long ix = i;")
        out.append("        (void)ix;")
        for i in used:
            out.append(f"        const {CTYPE[o.params[i]]} a{i} = s{i}[i];")
        out.append(f"        d[i] = ({rt})({_expr(o, o.body)});")
        out.append("    }")
        out.append("}")
        hparams = ["vec_p d"] + [f"vec_p s{i}" for i in range(len(o.params))] + ["int len", "vec_p scratch"]
        out.append(f'extern "C" void fused{k}({", ".join(hparams)}) {{')
        out.append("    (void)scratch;")
        passed = set(used) | set(gsrc) | set(gdesc)
        for i in range(len(o.params)):
            if i not in passed:
                out.append(f"    (void)s{i};")
        out.append("    if (len <= 0) return;")
        args = [f"({rt} *)d"] + [f"(const {CTYPE[o.params[i]]} *)s{i}" for i in used]
        args += [f"(const {CTYPE[o.params[i]]} *)s{i}" for i in sorted(gsrc) if i not in used]
        args += [f"(const int *)s{j}" for j in gdesc]
        args.append("(long long)len")
        out.append(f"    k_fused{k}<<<fused_grid(len), 256, 0, (hipStream_t)cvlhip_get_stream()>>>({', '.join(args)});")
        out.append(f'    fused_check("fused{k}");')
        out.append("}")
        out.append(f'extern "C" int fused{k}_scratch(int len) {{ (void)len; return 0; }}')
        # argument 0 may be overwritten unless a gather reads it elsewhere
        inplace = 1 if o.params and o.params[0] != "SEGDES" and 0 not in gsrc else 0
        out.append(f'extern "C" unsigned fused{k}_inplace(void) {{ return {inplace}u; }}')
        out.append("")
    # descriptors
    out.append('extern "C" {')
    out.append(f"int vops_count = {len(ops)};")
    out.append("vopdes_t vops[] = {")
    for o in ops:
        k, n = o.number, len(o.params)
        atypes = ", ".join(VTYPE[t] for t in o.params) or "Illegal"
        # every vector argument has the element length of argument 0; a
        # segment descriptor describes that many elements
        lens = []
        a0_seg = bool(o.params) and o.params[0] == "SEGDES"
        gsrc = S.gather_sources(o.body)
        gseg = {v: k for k, v in gsrc.items()}
        # the vectors of element length: the rules can only compare with
        # the first five arguments
        elem = [i for i, u in enumerate(o.params) if u != "SEGDES" and i < 5 and i not in gsrc]

        def describes_elements(j):
            # a descriptor over the elements: argument 1 itself defines the
            # length, another one must describe a vector of that length (with
            # none to compare with, the fuser's size analysis is the check)
            if j == 0:
                return [f"COMPAT{i + 1}" for i in elem]
            return [f"COMPAT{elem[0] + 1}"] if elem else []

        for j, t in enumerate(o.params):
            bits = []
            if j in gsrc:
                pass                                   # a gathered vector has its own length
            elif j in gseg:
                if gseg[j] < 5:
                    bits.append(f"COMPAT{gseg[j] + 1}")  # describes the gathered vector
            elif t == "SEGDES":
                bits += describes_elements(j)
            elif j > 0 and not a0_seg:
                bits.append("AGREE1")
            lens.append("|".join(bits) if bits else "NONE")
        rlen = "COMPAT1" if a0_seg else "AGREE1"
        scr = "1"
        if o.reduce:
            r = o.reduce[1]
            if r >= 5:
                raise S.DefinitionError("a reduced descriptor must be one of the first five arguments")
            lens[r] = "|".join(describes_elements(r)) or "NONE"   # describes the reduced elements
            rlen = f"AGREE{r + 1}"                     # one result per segment
            scr = str(r + 1)                           # scratch sized by (n, m) of the descriptor
        out.append(f'    {{VCODE_FUSED_TOKEN, "fused{k}", {n}, 1, {{{atypes}}}, {{{", ".join(lens) or "NONE"}}}, '
                   f'{{{VTYPE[o.result]}}}, {{{rlen}}}, {{{scr}}}, Elwise{n}}},')
    out.append("};")
    out.append("cvl_triple_t cvl_funs[] = {")
    for o in ops:
        k = o.number
        out.append(f"    {{{{(void (*)(void))fused{k}, (int (*)(void))fused{k}_scratch, "
                   f"(unsigned (*)(void))fused{k}_inplace}}, {{0, 0, 0}}, {{0, 0, 0}}}},")
    out.append("};")
    out.append("void init(void *mem) { (void)mem; }")
    out.append("}")
    return "\n".join(out) + "\n"


def build(defs_text: str, out_so: str, hip_path: str | None = None, verbose: bool = False) -> str:
    ops = S.parse(defs_text)
    name = os.path.splitext(os.path.basename(out_so))[0]
    src = generate(ops, name)
    hip_path = hip_path or os.path.splitext(out_so)[0] + ".hip"
    os.makedirs(os.path.dirname(os.path.abspath(out_so)), exist_ok=True)
    with open(hip_path, "w") as f:
        f.write(src)
    tuned = os.path.join(PKG, "csrc", "tuned.hpp")      # the library's tile shapes (same tiles, same bits)
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
           *(["-include", tuned] if os.path.exists(tuned) else []),
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "csrc"),
           hip_path, "-o", out_so + ".tmp",
           # the CVL library supplies the stream; found relative to the library's directory
           "-L", PKG, "-lcvl_hip",
           "-Wl,-rpath,$ORIGIN/" + os.path.relpath(PKG, os.path.dirname(os.path.abspath(out_so)))]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError(f"hipcc failed on {hip_path}")
    os.replace(out_so + ".tmp", out_so)
    if verbose:
        print(f"built {out_so} ({len(ops)} fused ops)")
    return out_so
