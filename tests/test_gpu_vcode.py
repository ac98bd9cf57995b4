"""Whole VCODE programs on the GPU: the interpreter over libcvl_hip.so must
print exactly what the reference serial CVL printed for the same program
(tests/golden/vcode, recorded by tests/golden/gen_vcode_golden.py), apart
from timer values.  This covers the interpreter's call sites (argument
order of every CVL call shape, result allocation, in-place reuse, scratch
sizing) and every CVL entry point the programs reach, including nesltest,
the reference's self-checking test program (its expected "failed" lines
are those the serial CVL prints too).
"""
import hashlib
import json
import os

import pytest

from neslgpu_amd import vinterp as V
from neslgpu_amd.build import FUSED_DIR

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "vcode")
PROGS = os.path.join(GOLD, "programs")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))


def stable(stdout, volatile):
    return [l for i, l in enumerate(stdout.splitlines()) if i not in volatile]


def variant_text(spec):
    text = V.read(os.path.join(PROGS, spec["base"]))
    if spec.get("full"):
        text = V.full_result(text)
    if spec.get("size"):
        text = V.set_size(text, spec["size"])
    if spec.get("key_range"):
        text = V.set_function_int_const(text, "QS_4", 100, spec["key_range"])
    return text


@pytest.mark.parametrize("name", sorted(k for k in MANIFEST if k.endswith(".vcode")))
def test_program_matches_serial_cvl(name):
    print("vcode program", name, flush=True)
    rec = MANIFEST[name]
    r = V.run(os.path.join(PROGS, name), timeout=240)
    want = open(os.path.join(GOLD, name + ".out")).read()
    assert stable(r.stdout, rec["volatile_stdout_lines"]) == stable(want, rec["volatile_stdout_lines"])
    assert r.stderr == open(os.path.join(GOLD, name + ".err")).read()


def fused_library(rec):
    path = os.path.join(FUSED_DIR, rec["superops"][: -len(".superops")] + ".so")
    assert os.path.exists(path), f"{path} not built (python -m neslgpu_amd.build)"
    return path


@pytest.mark.parametrize("name", sorted(k for k in MANIFEST if k.endswith(".fcode")))
def test_fused_program_matches_unfused_serial_cvl(name):
    """FUSED n instructions run kernels generated from the super-op
    definitions; the program must print what its unfused form printed."""
    print("fused program", name, flush=True)
    rec = MANIFEST[name]
    r = V.run(variant_text(rec), fused=fused_library(rec), timeout=240)
    want = open(os.path.join(GOLD, rec["expected"] + ".out")).read()
    assert stable(r.stdout, rec["volatile_stdout_lines"]) == stable(want, rec["volatile_stdout_lines"])
    assert r.stderr == open(os.path.join(GOLD, rec["expected"] + ".err")).read()


@pytest.mark.parametrize("name", sorted(k for k in MANIFEST if not k.endswith("code") and "own" not in MANIFEST[k]))
def test_program_variant_matches_serial_cvl(name):
    print("vcode variant", name, flush=True)
    rec = MANIFEST[name]
    fused = fused_library(rec) if "superops" in rec else None
    r = V.run(variant_text(rec), fused=fused, timeout=240)
    got = stable(r.stdout, rec["volatile_stdout_lines"])
    if rec.get("digest"):
        assert len(r.stdout.splitlines()) == rec["stdout_lines"]
        assert hashlib.sha256("\n".join(got).encode()).hexdigest() == rec["sha256"]
    else:
        want = stable(open(os.path.join(GOLD, rec.get("expected", name) + ".out")).read(),
                      rec["volatile_stdout_lines"])
        if "superops" in rec:
            # dropping the fused program's post-timer code can leave values
            # below the result that the unfused program never had; the timer
            # value is the last line
            got = r.stdout.splitlines()[:-1][-len(want):]
        assert got == want
