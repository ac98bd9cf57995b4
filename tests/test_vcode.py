"""The VCODE interpreter without a GPU.

- Every fixture program parses and links (-P).
- The program rewrites of neslgpu_amd.vinterp (size, key range, full
  result) do what they say.
- The recorded golden output of quicksort on 2^20 keys in [0, 2^31-1) is
  pinned by mathematics: it must be the sorted lrand48 stream that RAND
  produces from an unseeded state (oracle.Lrand48), not only what the serial
  CVL printed.
- With the reference serial CVL built (oracle/_ref), the interpreter over
  it reproduces every recorded output (the recording itself came from that
  pair, so this guards the interpreter against regressions).
"""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from neslgpu_amd import vinterp as V

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden", "vcode")
PROGS = os.path.join(GOLD, "programs")
SERIAL = os.path.join(ROOT, "oracle", "_ref", "libcvl_serial.so")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))


def variant_text(spec):
    text = V.read(os.path.join(PROGS, spec["base"]))
    if spec.get("full"):
        text = V.full_result(text)
    if spec.get("size"):
        text = V.set_size(text, spec["size"])
    if spec.get("key_range"):
        text = V.set_function_int_const(text, "QS_4", 100, spec["key_range"])
    return text


def stable(stdout, volatile):
    return [l for i, l in enumerate(stdout.splitlines()) if i not in volatile]


def digest(lines):
    return hashlib.sha256("\n".join(lines).encode()).hexdigest()


def need_vinterp():
    if not os.path.exists(V.VINTERP):
        pytest.skip("vinterp not built")


@pytest.mark.parametrize("name", sorted(os.listdir(PROGS)))
def test_programs_parse(name):
    need_vinterp()
    r = subprocess.run([V.VINTERP, "-P", os.path.join(PROGS, name)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "instructions" in r.stdout


def test_rewrites():
    text = V.read(os.path.join(PROGS, "qs.1000.vcode"))
    t2 = V.set_size(text, 12345)
    assert "CONST INT 12345" in t2.splitlines()[3]
    t3 = V.set_function_int_const(t2, "QS_4", 100, 2147483647)
    assert t3.count("CONST INT 2147483647") == 1
    full = V.full_result(text)
    lines = full.splitlines()
    i = lines.index("STOP_TIMER")
    assert lines[i + 1] == "RET"
    assert len(lines) == len(text.splitlines()) - 4


def test_qs_large_golden_is_the_sorted_rand_stream():
    """qs on 2^20 keys: output = [ n ] + the sorted keys RAND drew."""
    spec = MANIFEST["qs.1048576.full"]
    n = spec["size"]
    keys = __import__("oracle.oracle", fromlist=["Lrand48"]).Lrand48(0).rnd_wuz(
        np.full(n, spec["key_range"], np.int32))
    assert keys.min() >= 0 and keys.max() < spec["key_range"]
    lines = [f"[ {n} ]", "(" + "".join(f" {k}" for k in np.sort(keys)) + " )"]
    assert digest(lines) == spec["sha256"]


@pytest.mark.skipif(not os.path.exists(SERIAL), reason="reference serial CVL not built")
@pytest.mark.parametrize("name", sorted(k for k in MANIFEST if k.endswith(".vcode")))
def test_serial_reproduces_golden(name):
    need_vinterp()
    rec = MANIFEST[name]
    r = V.run(os.path.join(PROGS, name), backend=SERIAL, arena_units=200_000_000, timeout=300)
    want = open(os.path.join(GOLD, name + ".out")).read()
    assert stable(r.stdout, rec["volatile_stdout_lines"]) == stable(want, rec["volatile_stdout_lines"])
    err = "".join(l for l in r.stderr.splitlines(keepends=True) if not l.startswith("rnd_wuz:"))
    assert err == open(os.path.join(GOLD, name + ".err")).read()


def _serial_or_skip():
    if not os.path.exists(SERIAL):
        pytest.skip("reference serial CVL not built (make -C oracle)")


def test_diagnostic_switches(tmp_path):
    """The reference interpreter's debugging switches (nesl/vcode/main.c:
    130-139, 194-214) on quicksort of 1000 keys over the serial CVL: each
    trace prints what it names and the program's own output is unchanged;
    -p writes <program>.vprof in the reference's format
    (vcode-profile.c:104-131) with one line per function."""
    _serial_or_skip()
    src = tmp_path / "qs.1000.vcode"
    src.write_text(open(os.path.join(PROGS, "qs.1000.vcode")).read())
    want = open(os.path.join(GOLD, "qs.1000.vcode.out")).read()

    def run(*flags):
        r = subprocess.run([V.VINTERP, "-B", SERIAL, *flags, str(src)], capture_output=True, text=True,
                           timeout=120, cwd=tmp_path)
        assert r.returncode == 0, r.stderr[-500:]
        vol = MANIFEST["qs.1000.vcode"]["volatile_stdout_lines"]
        keep = lambda t: [l for i, l in enumerate(t.splitlines()) if i not in vol]
        assert keep(r.stdout) == keep(want), flags
        return r.stderr

    prog = V.read(str(src))
    funcs = [l.split()[1] for l in prog.splitlines() if l.startswith("FUNC ")]
    err = run("-d", "program")
    assert sorted(l.split()[1] for l in err.splitlines() if l.startswith("FUNC ")) == sorted(funcs)
    err = run("-d", "link")
    assert sum(l.startswith("link: ") for l in err.splitlines()) == prog.count("\nCALL ")
    err = run("-d", "lex")
    assert "lex: line 3 WORD FUNC" in err
    err = run("-d", "runtime")
    assert "call QUICKSORT_80" in err and "return from QUICKSORT_80" in err
    err = run("-d", "stack", "-d", "values", "-d", "heap")
    assert "stack after line 4 (1): INT[1]( 1000 )" in err and "heap after line 4:" in err
    run("-p", "-g", "-c")
    lines = (tmp_path / "qs.1000.vprof").read_text().splitlines()
    assert lines[0].split() == ["#", "Calls", "Time", "Min.", "Avg.", "Max.", "Cumulative", "Name"]
    rows = {l.split()[-1]: l.split() for l in lines[2:]}
    assert set(rows) == set(funcs)
    # calls counted; the entry's cumulative time holds its (non-recursive) callee's
    assert int(rows["QUICKSORT_80"][0]) > 1 and int(rows[funcs[0]][0]) == 1
    assert float(rows[funcs[0]][5]) >= float(rows["QS_4"][5])


def test_abort_switch(tmp_path):
    """-d abort turns an interpreter error into abort() (a core dump where
    the system allows one) instead of exit(1)."""
    _serial_or_skip()
    src = tmp_path / "bad.vcode"
    src.write_text("FUNC MAIN1_1\nPOP 1 0\nRET\n")
    r = subprocess.run([V.VINTERP, "-B", SERIAL, str(src)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "vinterp:" in r.stderr
    r = subprocess.run(["bash", "-c", f"ulimit -c 0; exec {V.VINTERP} -B {SERIAL} -d abort {src}"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode in (-6, 128 + 6), r.returncode      # SIGABRT
