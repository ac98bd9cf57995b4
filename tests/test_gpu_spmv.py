"""Nested SpMV (neslgpu_amd/vcode/programs/spmv.vcode) on the GPU, unfused,
fused by this package's fuser (neslgpu_amd/fusion/fuser.py: the gather of x
by column, the product with the values and the segmented sum over the rows
in one generated kernel) and with basic blocks replayed as HIP graphs (-G).

The serial reference CVL ran the same program (golden spmv.small.full):
row lengths, column indices, values and x are bit-exact; each y_i is a sum
whose association differs between a sequential fold and the GPU's, so y is
checked within 2 (ceil(log2 L) + 2) 2^-53 sum|x_c v| of the serial value
(L = row length).  Graph replays must reproduce the eager run bit for bit:
the program repeats the product K times and prints the last one, so K >= 2
prints a replayed block's result.
"""
import json
import math
import os

import numpy as np
import pytest

from neslgpu_amd import vinterp as V
from neslgpu_amd.build import generated

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "vcode")
PROG = os.path.join(os.path.dirname(HERE), "neslgpu_amd", "vcode", "programs")
REC = json.load(open(os.path.join(GOLD, "manifest.json")))["spmv.small.full"]


def parse_vec(line):
    body = line.strip()[1:-1].split()
    if line.lstrip().startswith("(") and any("e" in t or "." in t or "n" in t for t in body):
        return np.array([float(t) for t in body])
    return np.array([int(t) for t in body], np.int64)


def program(fused, k):
    text = V.read(generated("spmv", "fcode") if fused else os.path.join(PROG, "spmv.vcode"))
    text = V.set_function_int_const(text, "MAIN", 10, k)
    return V.full_result(text)


def run(fused, k, graphs):
    lib = generated("spmv", "so") if fused else None
    r = V.run(program(fused, k), fused=lib, graphs=graphs, stdin=REC["stdin"], timeout=120, timed=graphs)
    return r.stdout.splitlines()[:-1], r.stderr


@pytest.mark.parametrize("fused,graphs,k", [(False, False, 1), (True, False, 1), (False, True, 3),
                                             (True, True, 2), (True, True, 5)])
def test_spmv_matches_serial(fused, graphs, k):
    print("spmv fused", fused, "graphs", graphs, "k", k, flush=True)
    got, err = run(fused, k, graphs)
    want = open(os.path.join(GOLD, "spmv.small.full.out")).read().splitlines()[:-1]
    assert len(got) == len(want) == 9
    # L N Snnz Sx C V X R bit-exact
    assert got[:8] == want[:8]
    lens = parse_vec(want[0])
    cols = parse_vec(want[4])
    vals = parse_vec(want[5])
    x = parse_vec(want[6])
    y_ref = parse_vec(want[8])
    y = parse_vec(got[8])
    prod = np.abs(x[cols] * vals)
    absum = np.add.reduceat(np.append(prod, 0.0), np.concatenate([[0], np.cumsum(lens)[:-1]]))[: lens.size]
    absum[lens == 0] = 0.0
    tol = np.array([2 * (math.ceil(math.log2(max(int(L), 2))) + 2) * 2.0 ** -53 for L in lens]) * absum
    assert np.all(np.abs(y - y_ref) <= tol)
    if graphs:
        assert "captured" in err, err


def test_graph_replays_equal_eager():
    """Every replay count gives the bits of the eager run."""
    eager, _ = run(True, 1, False)
    for k in (2, 3, 4):
        got, err = run(True, k, True)
        assert got == eager, k
        assert "captured" in err


@pytest.mark.parametrize("name", ["qs.1000.vcode", "quickhull.1000.vcode", "bh.100.vcode", "nesltest.vcode"])
def test_graph_mode_does_not_change_programs(name):
    print("graph mode", name, flush=True)
    rec = json.load(open(os.path.join(GOLD, "manifest.json")))[name]
    r = V.run(os.path.join(GOLD, "programs", name), graphs=True, timeout=240)
    want = open(os.path.join(GOLD, name + ".out")).read().splitlines()
    got = r.stdout.splitlines()
    keep = [i for i in range(len(want)) if i not in rec["volatile_stdout_lines"]]
    assert [got[i] for i in keep] == [want[i] for i in keep]
