"""The SpMV workload's inputs and reference, pinned against what the
reference serial CVL printed.

workloads.spmv_inputs redraws the columns, values and x that the SpMV
program draws with RAND; on the small program (tests/golden/vcode/
spmv.small.full, run on the serial CVL) the columns must be equal and the
values, x and y equal to the 15 printed digits.  At the benchmark's size
(2^26 nonzeros) the serial CVL's y at the sampled rows
(spmv.4194304.sample.npz) must lie within the error bound of a sequential
double sum of the extended-precision row sums the GPU test checks against.
"""
import hashlib
import os

import numpy as np

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "vcode")
SMALL_ROWS = [3, 0, 2, 1, 17, 1, 1, 64, 0, 5] * 30 + [300, 2, 7]


def nums(line, dtype):
    return np.array(line.strip("()[] ").split(), dtype=dtype)


def test_spmv_inputs_match_the_serial_cvl_small():
    from neslgpu_amd import workloads as W
    out = open(os.path.join(GOLD, "spmv.small.full.out")).read().splitlines()
    rows = np.array(SMALL_ROWS, np.int32)
    cols, vals, x = W.spmv_inputs(rows, 257, O.Lrand48(0).rnd_wuz)
    assert np.array_equal(nums(out[4], np.int64), cols)
    assert np.all(np.abs(nums(out[5], np.float64) - vals) <= 5e-15 * np.abs(vals) + 1e-300)
    assert np.all(np.abs(nums(out[6], np.float64) - x) <= 5e-15 * np.abs(x) + 1e-300)
    y, ab = W.spmv_reference(rows, cols, vals, x)
    yp = nums(out[8], np.float64)
    seq = (np.maximum(rows, 1) - 1) * 2.0 ** -53 * ab
    assert np.all(np.abs(yp - y.astype(np.float64)) <= seq + 5e-15 * np.abs(yp) + 1e-300)


def test_spmv_y_program_rewrite():
    from neslgpu_amd import workloads as W
    text = open(os.path.join(ROOT, "neslgpu_amd", "vcode", "programs", "spmv.vcode")).read()
    t = W.spmv_y_program(text)
    assert "CONST INT 1\nSTART_TIMER" in t and "+_REDUCE FLOAT\nCPOP 1 1\nRET" not in t


def test_full_size_serial_sample_against_extended_sum():
    from neslgpu_amd import workloads as W
    rows = W.spmv_rows()
    assert int(rows.sum()) == W.SPMV_NNZ and rows.size == W.SPMV_ROWS and int(rows.max()) <= W.SPMV_LMAX
    with np.load(os.path.join(GOLD, "spmv.4194304.sample.npz"), allow_pickle=False) as z:
        assert str(z["rows_sha256"]) == hashlib.sha256(rows.tobytes()).hexdigest()
        idx, ys, lens = z["rows_index"], z["y"], z["row_len"]
    assert np.array_equal(rows[idx], lens)
    cols, vals, x = W.spmv_inputs(rows, W.SPMV_ROWS, O.Lrand48(0).rnd_wuz)
    y, ab = W.spmv_reference(rows, cols, vals, x)
    seq = (np.maximum(rows[idx].astype(np.int64), 1) - 1) * 2.0 ** -53 * ab[idx]
    err = np.abs(ys - y[idx].astype(np.float64))
    assert np.all(err <= seq + 5e-15 * np.abs(ys) + 1e-300)
    # the exact sums are not all trivially zero, and the long rows are in the sample
    assert np.count_nonzero(ys) > idx.size // 2 and int(lens.max()) == int(rows.max())
