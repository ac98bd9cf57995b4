"""Golden outputs of whole VCODE programs, produced by the serial reference CVL.

Runs this package's VCODE interpreter (neslgpu_amd/bin/vinterp) with the
reference serial CVL compiled from its sources (oracle/_ref/libcvl_serial.so,
built by oracle/Makefile) as the backend, on VCODE programs that the
reference ships as benchmark and test inputs, and records what each program
prints:

    tests/golden/vcode/programs/<name>      the program (input data, copied
                                            from the reference's bench/NESL and
                                            nesl/vcode directories)
    tests/golden/vcode/<name>.out           stdout (the final stack)
    tests/golden/vcode/<name>.err           stderr (WRITE output)
    tests/golden/vcode/manifest.json        per program: stdout lines that hold
                                            timings (they differ run to run)

The serial CVL's rnd_wuz prints a diagnostic line ("rnd_wuz: ...") on each
call; those lines are dropped from the recorded stderr.  Fused programs
(.fcode) compute the same values as their unfused form, so their expected
output is the unfused program's.

    python tests/golden/gen_vcode_golden.py
"""
import hashlib
import json
import tempfile
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
OUT = os.path.join(HERE, "vcode")
PROGS = os.path.join(OUT, "programs")
VINTERP = os.path.join(ROOT, "neslgpu_amd", "bin", "vinterp")
SERIAL = os.path.join(ROOT, "oracle", "_ref", "libcvl_serial.so")

# (fixture name, path in the reference)
PROGRAMS = [
    ("nesltest.vcode", "nesl/vcode/nesltest.vcode"),
    ("qs.1000.vcode", "bench/NESL/qs.1000.vcode"),
    ("quickhull.1000.vcode", "bench/NESL/quickhull.1000.vcode"),
    ("dotp.1000.vcode", "bench/NESL/dotp.1000.vcode"),
    ("black-scholes.1000.vcode", "bench/NESL/black-scholes.1000.vcode"),
    ("bh.100.vcode", "bench/NESL/bh.100.vcode"),
    ("radix.1000000.vcode", "bench/NESL/radix.1000000.vcode"),
    ("delaunay.10000.vcode", "bench/NESL/delaunay.10000.vcode"),
    ("muladd.vcode", "bench/NESL/muladd.vcode"),
    # fused forms (FUSED n calls into a library generated from the super-op
    # definitions in tests/golden/vcode/superops/)
    ("qs.1000.fcode", "bench/NESL/qs.1000.fcode"),
    ("quickhull.1000.fcode", "bench/NESL/quickhull.1000.fcode"),
    ("dotp.1000.fcode", "bench/NESL/dotp.1000.fcode"),
    ("muladd.fcode", "bench/NESL/muladd.fcode"),
    ("bh.100.fcode", "bench/NESL/bh.100.fcode"),
]

# super-op definitions: the "fused OPk#c (...) = ..." lines that the fusion
# compiler wrote at the end of each generated library source
SUPEROPS = [
    ("qs.1000", "bench/NESL/qs.1000.cu"),
    ("quickhull.1000", "bench/NESL/quickhull.1000.cu"),
    ("dotp.1000", "bench/NESL/dotp.1000.cu"),
    ("muladd", "bench/NESL/muladd.cu"),
    ("bh.100", "bench/NESL/bh.100.cu"),
]

# variants of the benchmark programs, built by neslgpu_amd.vinterp's rewrites:
# the whole result printed (full_result), other sizes (set_size), and for
# quicksort the key range of the config the benchmark is quoted on (keys
# uniform in [0, 2^31-1)).  Large outputs are recorded as a SHA-256 digest.
VARIANTS = [
    ("qs.1000.full", {"base": "qs.1000.vcode", "full": True}),
    ("quickhull.1000.full", {"base": "quickhull.1000.vcode", "full": True}),
    ("qs.1048576.full", {"base": "qs.1000.vcode", "full": True, "size": 1 << 20,
                         "key_range": 2147483647, "digest": True}),
    ("quickhull.100000.full", {"base": "quickhull.1000.vcode", "full": True, "size": 100000, "digest": True}),
    # the benchmark's size: 2^24 points in the unit disk, the whole hull
    ("quickhull.16777216.full", {"base": "quickhull.1000.vcode", "full": True, "size": 1 << 24,
                                 "digest": True}),
    # fused programs print what their unfused forms print
    ("qs.1000.fcode.full", {"base": "qs.1000.fcode", "full": True, "same_as": "qs.1000.full",
                            "superops": "qs.1000.superops"}),
    ("quickhull.1000.fcode.full", {"base": "quickhull.1000.fcode", "full": True,
                                   "same_as": "quickhull.1000.full", "superops": "quickhull.1000.superops"}),
]

# this package's own programs (neslgpu_amd/vcode/programs), with their stdin
OWN = os.path.join(ROOT, "neslgpu_amd", "vcode", "programs")
SPMV_ROWS = [3, 0, 2, 1, 17, 1, 1, 64, 0, 5] * 30 + [300, 2, 7]
OWN_VARIANTS = [
    ("spmv.small.full", {"own": "spmv.vcode", "full": True,
                         "stdin": "( " + " ".join(map(str, SPMV_ROWS)) + " )\n( 257 )\n"}),
]

ARENA_UNITS = "200000000"

# SpMV at the benchmark's size (2^22 rows, power-law lengths with mean 16
# and maximum 2^16, 2^26 nonzeros), one product, y printed: the serial CVL's
# y at a sample of rows (every 1024th row and the 64 longest rows) is kept
# as a fixture; the whole y is checked against an extended-precision sum.
SPMV_SAMPLE = os.path.join(OUT, "spmv.4194304.sample.npz")


def spmv_full_sample():
    sys.path.insert(0, ROOT)
    import numpy as np
    from neslgpu_amd import workloads as W
    rows = W.spmv_rows()
    text = W.spmv_y_program(open(os.path.join(OWN, "spmv.vcode")).read())
    with tempfile.NamedTemporaryFile("w", suffix=".vcode", delete=False) as f:
        f.write(text)
    r = subprocess.run([VINTERP, "-B", SERIAL, "-m", str((1 << 31) - 1), f.name], capture_output=True, text=True,
                       timeout=1200, input=W.spmv_stdin(rows, W.SPMV_ROWS))
    os.unlink(f.name)
    if r.returncode != 0:
        raise SystemExit(f"spmv: exit {r.returncode}\n{r.stderr[-2000:]}")
    lines = r.stdout.splitlines()
    y = np.array(lines[-2].strip("() ").split(), dtype=np.float64)
    assert y.size == rows.size, (y.size, rows.size)
    idx = np.union1d(np.arange(0, rows.size, 1024), np.argsort(rows, kind="stable")[-64:]).astype(np.int64)
    np.savez(SPMV_SAMPLE, rows_index=idx, y=y[idx], row_len=rows[idx],
             rows_sha256=np.array(hashlib.sha256(rows.tobytes()).hexdigest()))
    print("spmv full-size sample", idx.size, "rows", flush=True)


def variant_text(spec):
    """The program text of a variant (same rewrite the tests apply)."""
    sys.path.insert(0, ROOT)
    from neslgpu_amd import vinterp as V
    text = V.read(os.path.join(PROGS, spec["base"]))
    if spec.get("full"):
        text = V.full_result(text)
    if spec.get("size"):
        text = V.set_size(text, spec["size"])
    if spec.get("key_range"):
        text = V.set_function_int_const(text, "QS_4", 100, spec["key_range"])
    return text


def stable_digest(stdout, volatile):
    lines = [l for i, l in enumerate(stdout.splitlines()) if i not in volatile]
    return hashlib.sha256("\n".join(lines).encode()).hexdigest()


def masked(stdout, volatile):
    """stdout with the run-dependent (timer) lines replaced by a marker."""
    return "".join(("<timer>\n" if i in volatile else l + "\n") for i, l in enumerate(stdout.splitlines()))


def run(prog, stdin=None):
    r = subprocess.run([VINTERP, "-B", SERIAL, "-m", ARENA_UNITS, prog], capture_output=True, text=True,
                       timeout=600, input=stdin)
    if r.returncode != 0:
        raise SystemExit(f"{prog}: exit {r.returncode}\n{r.stderr[-2000:]}")
    err = "".join(l for l in r.stderr.splitlines(keepends=True) if not l.startswith("rnd_wuz:"))
    return r.stdout, err


def main():
    sys.path.insert(0, ROOT)
    from neslgpu_amd import vinterp as V
    if not os.path.exists(SERIAL):
        raise SystemExit("build the reference serial CVL first: make -C oracle")
    os.makedirs(PROGS, exist_ok=True)
    os.makedirs(os.path.join(OUT, "superops"), exist_ok=True)
    for name, rel in PROGRAMS:
        shutil.copyfile(os.path.join(REF, rel), os.path.join(PROGS, name))
    for name, rel in SUPEROPS:
        # a definition may wrap over several lines; it ends where the next
        # one starts or the comment block closes
        defs = []
        for l in open(os.path.join(REF, rel)):
            if l.startswith("fused OP"):
                defs.append(l.strip())
            elif defs and not l.startswith("*/") and defs[-1] is not None:
                defs[-1] += " " + l.strip()
            if l.startswith("*/") and defs:
                defs.append(None)
        with open(os.path.join(OUT, "superops", name + ".superops"), "w") as f:
            f.writelines(" ".join(d.split()) + "\n" for d in defs if d)
    manifest = {}
    for name, _ in PROGRAMS:
        if name.endswith(".fcode"):
            continue
        path = os.path.join(PROGS, name)
        out1, err1 = run(path)
        out2, err2 = run(path)
        l1, l2 = out1.splitlines(), out2.splitlines()
        assert len(l1) == len(l2) and err1 == err2, name
        volatile = [i for i, (a, b) in enumerate(zip(l1, l2)) if a != b]
        # the last stdout line of a benchmark program is its STOP_TIMER value
        if name != "nesltest.vcode" and l1 and l1[-1].startswith("( ") and "e" in l1[-1] and \
                (len(l1) - 1) not in volatile:
            volatile.append(len(l1) - 1)
        with open(os.path.join(OUT, name + ".out"), "w") as f:
            f.write(masked(out1, volatile))
        with open(os.path.join(OUT, name + ".err"), "w") as f:
            f.write(err1)
        manifest[name] = {"volatile_stdout_lines": volatile, "stdout_lines": len(l1)}
        print(name, len(l1), "lines, volatile", volatile, flush=True)
    for name, spec in VARIANTS:
        if "same_as" in spec:
            continue
        with tempfile.NamedTemporaryFile("w", suffix=".vcode", delete=False) as f:
            f.write(variant_text(spec))
        out1, err1 = run(f.name)
        os.unlink(f.name)
        l1 = out1.splitlines()
        volatile = [len(l1) - 1]                 # the timer value
        rec = dict(spec, volatile_stdout_lines=volatile, stdout_lines=len(l1))
        if spec.get("digest"):
            rec["sha256"] = stable_digest(out1, volatile)
        else:
            with open(os.path.join(OUT, name + ".out"), "w") as g:
                g.write(masked(out1, volatile))
        manifest[name] = rec
        print(name, len(l1), "lines", flush=True)
    for name, spec in OWN_VARIANTS:
        text = open(os.path.join(OWN, spec["own"])).read()
        if spec.get("full"):
            text = V.full_result(text)
        with tempfile.NamedTemporaryFile("w", suffix=".vcode", delete=False) as f:
            f.write(text)
        out1, _ = run(f.name, spec["stdin"])
        os.unlink(f.name)
        l1 = out1.splitlines()
        volatile = [len(l1) - 1]
        with open(os.path.join(OUT, name + ".out"), "w") as g:
            g.write(masked(out1, volatile))
        manifest[name] = dict(spec, volatile_stdout_lines=volatile, stdout_lines=len(l1))
        print(name, len(l1), "lines", flush=True)
    for name, spec in VARIANTS:
        if "same_as" in spec:
            manifest[name] = dict(manifest[spec["same_as"]], **spec)
            manifest[name]["expected"] = spec["same_as"]
    for name, _ in PROGRAMS:
        if name.endswith(".fcode"):
            base = name[: -len(".fcode")] + ".vcode"
            # a fused program runs at its unfused form's size (some were
            # generated for another size than their file name says)
            size = int(V.entry_size(V.read(os.path.join(PROGS, base))))
            manifest[name] = dict(manifest[base], expected=base, superops=name[: -len(".fcode")] + ".superops",
                                  base=name, size=size)
    for name, _ in PROGRAMS:
        manifest.setdefault(name, {})["sha256_of_program"] = hashlib.sha256(
            open(os.path.join(PROGS, name), "rb").read()).hexdigest()
        manifest[name]["source"] = dict(PROGRAMS)[name]
    with open(os.path.join(OUT, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    spmv_full_sample()


if __name__ == "__main__":
    sys.exit(main())
