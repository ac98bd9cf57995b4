"""Synthetic inputs of the benchmark workloads, shared by bench.py and the
tests that check those workloads at full size.

  geometric_lengths   segment lengths >= 1, geometric, summing to n
  power_law_rows      CSR row lengths from a discrete power law with a given
                      mean and maximum, summing to exactly nnz
  spmv_stdin          the text the SpMV program READs: ( row lengths ) ( ncols )
  spmv_y_program      the SpMV program rewritten to run one product and leave
                      y (instead of sum(y)) on the stack, so it is printed
  spmv_inputs         the program's own random inputs (column indices,
                      values, x), drawn here from the same lrand48 stream
                      the program's RAND instructions draw from
"""
from __future__ import annotations

import numpy as np

SEED = 0x4E45534C00000000

# config 4: CSR 2^22 rows, power-law lengths (mean 16, max 2^16), 2^26 nnz
SPMV_ROWS = 1 << 22
SPMV_NNZ = 1 << 26
SPMV_MEAN = 16.0
SPMV_LMAX = 1 << 16
SPMV_SEED = SEED + 4 & 0x7FFFFFFF


def geometric_lengths(n: int, mean: int, seed: int) -> np.ndarray:
    """Segment lengths >= 1 with a geometric distribution, summing to n."""
    rng = np.random.default_rng(seed)
    m_guess = n // mean + 1024
    lens = rng.geometric(1.0 / mean, size=m_guess).astype(np.int64)
    cs = np.cumsum(lens)
    k = int(np.searchsorted(cs, n))
    lens = lens[: k + 1]
    lens[-1] -= int(cs[k] - n)
    lens = lens[lens > 0]
    assert int(lens.sum()) == n
    return lens.astype(np.int32)


def power_law_rows(m: int, nnz: int, mean: float, lmax: int, seed: int) -> np.ndarray:
    """m row lengths in [0, lmax] from a discrete power law with the given
    mean, adjusted by +-1 steps so they sum to exactly nnz."""
    L = np.arange(1, lmax + 1, dtype=np.float64)
    lo, hi = 1.01, 4.0
    for _ in range(60):                       # exponent giving the mean
        a = 0.5 * (lo + hi)
        p = L ** -a
        if (p * L).sum() / p.sum() > mean:
            lo = a
        else:
            hi = a
    p = L ** -a
    rng = np.random.default_rng(seed)
    rows = rng.choice(np.arange(1, lmax + 1), size=m, p=p / p.sum()).astype(np.int64)
    diff = nnz - int(rows.sum())
    while diff != 0:
        k = min(abs(diff), m)
        idx = rng.choice(m, size=k, replace=False)
        if diff > 0:
            idx = idx[rows[idx] < lmax]
            rows[idx] += 1
            diff -= idx.size
        else:
            idx = idx[rows[idx] > 0]
            rows[idx] -= 1
            diff += idx.size
    return rows.astype(np.int32)


def spmv_rows() -> np.ndarray:
    return power_law_rows(SPMV_ROWS, SPMV_NNZ, SPMV_MEAN, SPMV_LMAX, SPMV_SEED)


def spmv_stdin(rows: np.ndarray, ncols: int) -> str:
    return "( " + " ".join(map(str, np.asarray(rows).tolist())) + " )\n( " + str(int(ncols)) + " )\n"


# MAIN of spmv.vcode after STOP_TIMER: drop the 8 inputs below (y, timer),
# then sum y; the variant keeps y.
_SUM_TAIL = "CPOP 1 1\nCOPY 1 0\nLENGTH FLOAT\nMAKE_SEGDES\n+_REDUCE FLOAT\nCPOP 1 1\nRET"


def spmv_y_program(text: str) -> str:
    """spmv.vcode with one product (REPEAT count 1) whose result vector y is
    left on the stack under the timer value, so vinterp prints y."""
    t = text.replace("CONST INT 10\nSTART_TIMER", "CONST INT 1\nSTART_TIMER", 1)
    if t == text or _SUM_TAIL not in t:
        raise ValueError("not the SpMV program this rewrite expects")
    return t.replace(_SUM_TAIL, "RET", 1)


QS_KEYS = 1 << 20


def qs_bench_program(text: str) -> str:
    """qs.1000.vcode as the benchmark runs it: 2^20 keys drawn uniform in
    [0, 2^31 - 1) (the size constant of the entry function and QS_4's RAND
    range, 100 in the shipped program)."""
    from . import vinterp as V
    return V.set_function_int_const(V.set_size(text, QS_KEYS), "QS_4", 100, 2147483647)


def qs_bench_full_program(text: str) -> str:
    """The same, printing the sorted keys (the interpreter's full result)."""
    from . import vinterp as V
    return V.full_result(qs_bench_program(text))


def spmv_inputs(rows: np.ndarray, ncols: int, rnd_wuz):
    """(cols, vals, x) exactly as the SpMV program draws them with RAND:
    cols = RAND(ncols) per nonzero, then vals = RAND(2^30) / 2^29 - 1 per
    nonzero, then x = RAND(2^30) / 2^29 - 1 per column.  `rnd_wuz(s)` is a
    lrand48 stream's RAND (d[k] = lrand48() % s[k], in order)."""
    nnz = int(np.asarray(rows, np.int64).sum())
    cols = rnd_wuz(np.full(nnz, ncols, np.int32))
    vals = rnd_wuz(np.full(nnz, 1 << 30, np.int32)).astype(np.float64) / 536870912.0 - 1.0
    x = rnd_wuz(np.full(ncols, 1 << 30, np.int32)).astype(np.float64) / 536870912.0 - 1.0
    return cols, vals, x


def spmv_reference(rows: np.ndarray, cols, vals, x):
    """(y, bound): y = A x per row, summed in extended precision
    (numpy longdouble, 64-bit significand on x86-64), and sum |a_ij x_j|
    per row for the error bound of a double-precision sum."""
    rows = np.asarray(rows, np.int64)
    prod = vals * x[cols]                     # the double products the program forms
    starts = np.concatenate([[0], np.cumsum(rows)[:-1]])
    nz = rows > 0
    y = np.zeros(rows.size, np.longdouble)
    ab = np.zeros(rows.size, np.float64)
    if nz.any():
        y[nz] = np.add.reduceat(prod.astype(np.longdouble), starts[nz])
        ab[nz] = np.add.reduceat(np.abs(prod), starts[nz])
    return y, ab
