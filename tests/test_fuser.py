"""The fusion front end (neslgpu_amd/fusion/fuser.py) on the reference's own
VCODE programs, on the host.

Each program is fused; its super-op definitions must parse and type-check;
the fused program, with every FUSED n written back as the VCODE instructions
of super-op n's tree (tests/fusion_expand.py), must print on the reference
serial CVL exactly what the unfused program printed there (the goldens of
tests/golden/vcode, made by gen_vcode_golden.py).  That checks the
transformation itself -- regions, stack layouts, trees, argument order --
independently of the HIP code generated from the definitions, which
test_gpu_fuser.py runs.  The SpMV program must fuse into a +_REDUCE root
over a product with a gather of x.
"""
import json
import os

import pytest

import fusion_expand as X
from neslgpu_amd.fusion import fuser as F
from neslgpu_amd.fusion import superops as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "vcode")
PROGS = os.path.join(GOLD, "programs")
OWN = os.path.join(ROOT, "neslgpu_amd", "vcode", "programs")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))
SERIAL = os.path.join(ROOT, "oracle", "_ref", "libcvl_serial.so")


def stable(text, volatile):
    return [l for i, l in enumerate(text.splitlines()) if i not in volatile]


def serial_run(text, stdin=None):
    from neslgpu_amd import vinterp as V
    if not os.path.exists(SERIAL):
        pytest.skip("reference serial CVL not built (make -C oracle)")
    r = V.run(text, backend=SERIAL, timeout=300, stdin=stdin)
    err = "".join(l for l in r.stderr.splitlines(keepends=True) if not l.startswith("rnd_wuz:"))
    return r.stdout, err


def fused_and_expanded(text):
    fcode, sops = F.fuse_program(text)
    ops = S.parse(sops)
    assert all(o.result for o in ops)
    return X.expand_program(fcode, sops), fcode, ops


@pytest.mark.parametrize("name", sorted(k for k in MANIFEST if k.endswith(".vcode")))
def test_fused_program_prints_what_the_serial_cvl_printed(name):
    rec = MANIFEST[name]
    text = open(os.path.join(PROGS, name)).read()
    expanded, fcode, ops = fused_and_expanded(text)
    assert ops and "FUSED" in fcode, "nothing fused"
    out, err = serial_run(expanded)
    want = open(os.path.join(GOLD, name + ".out")).read()
    assert stable(out, rec["volatile_stdout_lines"]) == stable(want, rec["volatile_stdout_lines"])
    assert err == open(os.path.join(GOLD, name + ".err")).read()


@pytest.mark.parametrize("name", ["qs.1000.full", "quickhull.1000.full", "spmv.small.full"])
def test_fused_variants(name):
    from neslgpu_amd import vinterp as V
    rec = MANIFEST[name]
    if "own" in rec:
        text = open(os.path.join(OWN, rec["own"])).read()
    else:
        text = open(os.path.join(PROGS, rec["base"])).read()
    text = V.full_result(text)
    expanded, _, _ = fused_and_expanded(text)
    out, _ = serial_run(expanded, rec.get("stdin"))
    want = open(os.path.join(GOLD, name + ".out")).read()
    assert stable(out, rec["volatile_stdout_lines"]) == stable(want, rec["volatile_stdout_lines"])


def test_spmv_fuses_gather_product_and_row_sum():
    _, ops = F.fuse_program(open(os.path.join(OWN, "spmv.vcode")).read())
    roots = [o for o in S.parse(ops) if o.reduce and o.reduce[0] == "+_REDUCE"]
    assert len(roots) == 1
    body = roots[0].body
    assert body.op == "*" and any(isinstance(a, S.Apply) and a.op == "GATHER" and len(a.args) == 4
                                  for a in body.args)


def test_parse_emit_round_trip():
    text = open(os.path.join(PROGS, "nesltest.vcode")).read()
    funcs = F.parse(text)
    again = F.parse(F.emit(funcs))
    assert [(f.name, [x.text() for x in f.body]) for f in funcs] == \
        [(f.name, [x.text() for x in f.body]) for f in again]


def test_no_fusion_across_timers():
    """Regions end at START_TIMER and STOP_TIMER (fuse.sml:176-179): between
    the timers of the fused quicksort there is the timed call and stack
    moves, and nothing computed before the first timer moved past it."""
    fcode, _ = F.fuse_program(open(os.path.join(PROGS, "qs.1000.vcode")).read())
    lines = [l for l in fcode.splitlines() if l]
    i, j = lines.index("START_TIMER"), lines.index("STOP_TIMER")
    assert any(l.startswith("CALL ") for l in lines[i + 1:j])
    assert all(l.split()[0] in ("CALL", "COPY", "POP", "CPOP") for l in lines[i + 1:j])
