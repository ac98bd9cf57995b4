"""The SpMV program at the benchmark's size on the GPU (config 4: CSR 2^22
rows, power-law lengths with mean 16 and maximum 2^16, 2^26 nonzeros,
float64), through the interpreter on libcvl_hip.so: unfused (BPERMUTE, *,
+_REDUCE), fused (gather, multiply and segmented sum in one generated
kernel) and fused with the block replayed as a HIP graph.

The fused forms are the fuser's (neslgpu_amd/fusion/fuser.py) compilation of
the same program, built with the library.  The program draws its columns, values and x with RAND from the lrand48
stream; the test draws the same numbers with the oracle's lrand48
(workloads.spmv_inputs, pinned against the reference serial CVL's printed
inputs by tests/test_workloads.py) and sums every row in extended
precision.  Every printed y must lie within
    2 (ceil(log2 L) + 2) 2^-53 sum_j |a_ij x_j|
of that sum (plus the rounding of the 15-digit %.14e print), and within the
serial CVL's own error bound of the serial CVL's y at the rows kept in
tests/golden/vcode/spmv.4194304.sample.npz.  (Quickhull at 2^24 points is
checked bit-exact by digest in test_gpu_vcode.py.)"""
import hashlib
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAMPLE = os.path.join(ROOT, "tests", "golden", "vcode", "spmv.4194304.sample.npz")


@pytest.fixture(scope="module")
def reference():
    from neslgpu_amd import workloads as W
    rows = W.spmv_rows()
    cols, vals, x = W.spmv_inputs(rows, W.SPMV_ROWS, O.Lrand48(0).rnd_wuz)
    y, ab = W.spmv_reference(rows, cols, vals, x)
    return rows, y, ab


def printed_vector(line: str) -> np.ndarray:
    return np.array(line.strip("() ").split(), dtype=np.float64)


@pytest.mark.parametrize("form", ["unfused", "fused", "fused_graph"])
def test_spmv_full_size(reference, form):
    from neslgpu_amd import vinterp as V
    from neslgpu_amd import workloads as W
    from neslgpu_amd.build import generated
    rows, y_ref, ab = reference
    if form == "unfused":
        text = W.spmv_y_program(open(os.path.join(ROOT, "neslgpu_amd", "vcode", "programs", "spmv.vcode")).read())
        fused = None
    else:                           # the same variant, compiled by the fuser at build time
        text = open(generated("spmv_y", "fcode")).read()
        fused = generated("spmv_y", "so")
    print("spmv full size", form, rows.size, int(rows.sum()), flush=True)
    r = V.run(text, fused=fused, graphs=form == "fused_graph", stdin=W.spmv_stdin(rows, W.SPMV_ROWS),
              arena_units=(1 << 31) - 1, timeout=240)
    lines = r.stdout.splitlines()
    y = printed_vector(lines[-2])
    assert y.size == rows.size
    L = np.maximum(rows.astype(np.int64), 2)
    bound = 2.0 * (np.ceil(np.log2(L)) + 2) * 2.0 ** -53 * ab
    printing = 5e-15 * np.abs(y) + 1e-300              # %.14e keeps 15 significant digits
    err = np.abs(y - y_ref.astype(np.float64))
    bad = np.nonzero(err > bound + printing)[0]
    assert bad.size == 0, (form, bad[:5], err[bad[:5]], bound[bad[:5]])
    # the serial reference CVL's y at the sampled rows: both within their
    # own error bounds of the exact sum (the serial sum is sequential)
    with np.load(SAMPLE, allow_pickle=False) as z:
        assert str(z["rows_sha256"]) == hashlib.sha256(rows.tobytes()).hexdigest()
        idx, ys = z["rows_index"], z["y"]
    seq = (np.maximum(rows[idx].astype(np.int64), 1) - 1) * 2.0 ** -53 * ab[idx]
    tol = bound[idx] + seq + printing[idx] + 5e-15 * np.abs(ys)
    assert np.all(np.abs(y[idx] - ys) <= tol), form
