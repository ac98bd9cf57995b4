"""Python driver of the VCODE interpreter (neslgpu_amd/bin/vinterp).

    from neslgpu_amd import vinterp
    r = vinterp.run("qs.1000.vcode")                     # HIP CVL backend
    r = vinterp.run(text, backend="/path/libcvl_x.so")  # any CVL library
    print(r.stdout)

The interpreter is a native program (vcode/*.cpp) that loads the CVL
backend with dlopen; by default the backend is libcvl_hip.so of this
package, so vector instructions execute as HIP kernels on the GPU.  The
helpers below rewrite benchmark programs the way their sizes are varied:
set_size() changes the problem size constant of the entry function and
full_result() makes a benchmark print its whole result instead of one
element of it.
"""
from __future__ import annotations

import os
import re
import subprocess
import tempfile
from dataclasses import dataclass

HERE = os.path.dirname(os.path.abspath(__file__))
VINTERP = os.path.join(HERE, "bin", "vinterp")
HIP_BACKEND = os.path.join(HERE, "libcvl_hip.so")


@dataclass
class Result:
    returncode: int
    stdout: str
    stderr: str

    def values(self):
        """stdout lines (one per value left on the stack)."""
        return self.stdout.splitlines()


def run(program: str, *, backend: str | None = None, fused: str | None = None, arena_units: int | None = None,
        timeout: float | None = None, timed: bool = False, graphs: bool = False, check: bool = True,
        stdin: str | None = None) -> Result:
    """Run a VCODE program given as a path or as text."""
    if not os.path.exists(VINTERP):
        raise FileNotFoundError(f"{VINTERP} not built (python -m neslgpu_amd.build)")
    path, tmp = program, None
    if "\n" in program or not os.path.exists(program):
        tmp = tempfile.NamedTemporaryFile("w", suffix=".vcode", delete=False)
        tmp.write(program)
        tmp.close()
        path = tmp.name
    cmd = [VINTERP, "-B", backend or HIP_BACKEND]
    if fused:
        cmd += ["-l", fused]
    if arena_units:
        cmd += ["-m", str(int(arena_units))]
    if timed:
        cmd += ["-t"]
    if graphs:
        cmd += ["-G"]
    cmd.append(path)
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, input=stdin)
    finally:
        if tmp is not None:
            os.unlink(tmp.name)
    res = Result(r.returncode, r.stdout, r.stderr)
    if check and r.returncode != 0:
        raise RuntimeError(f"vinterp exited {r.returncode}: {r.stderr[-2000:]}")
    return res


def read(path: str) -> str:
    with open(path) as f:
        return f.read()


def _functions(lines):
    """[(name, first line index, RET line index)] of a VCODE text."""
    out, cur, start = [], None, 0
    for i, l in enumerate(lines):
        s = l.strip()
        if s.startswith("FUNC "):
            cur, start = s.split()[1], i
        elif s == "RET" and cur is not None:
            out.append((cur, start, i))
            cur = None
    return out


def entry_size(text: str) -> int:
    """The size constant (the first CONST INT) of the entry function."""
    lines = text.splitlines()
    name, start, end = _functions(lines)[0]
    for i in range(start + 1, end):
        m = re.match(r"^\s*CONST INT (-?\d+)\s*$", lines[i])
        if m:
            return int(m.group(1))
    raise ValueError(f"no size constant in {name}")


def set_size(text: str, n: int) -> str:
    """Replace the size constant (the first CONST INT) of the entry function."""
    lines = text.splitlines(keepends=True)
    name, start, end = _functions(lines)[0]
    for i in range(start + 1, end):
        if re.match(r"^\s*CONST INT -?\d+\s*$", lines[i]):
            lines[i] = f"CONST INT {int(n)}\n"
            return "".join(lines)
    raise ValueError(f"no size constant in {name}")


def set_function_int_const(text: str, func: str, old: int, new: int) -> str:
    """Replace the first `CONST INT old` inside function `func`."""
    lines = text.splitlines(keepends=True)
    for name, start, end in _functions(lines):
        if name != func:
            continue
        for i in range(start + 1, end):
            if lines[i].strip() == f"CONST INT {old}":
                lines[i] = f"CONST INT {new}\n"
                return "".join(lines)
    raise ValueError(f"CONST INT {old} not found in {func}")


def full_result(text: str) -> str:
    """Drop what a benchmark does between STOP_TIMER and its RET (picking one
    element of the result), so the whole result stays on the stack and is
    printed, followed by the timer value."""
    lines = text.splitlines(keepends=True)
    for name, start, end in _functions(lines):
        stops = [i for i in range(start, end) if lines[i].strip() == "STOP_TIMER"]
        if stops:
            s = stops[-1]
            return "".join(lines[: s + 1] + lines[end:])
    raise ValueError("no STOP_TIMER in the program")
