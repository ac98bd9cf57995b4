"""Shared helpers for the GPU parity tests: device tensors in, numpy out,
every call through the C-ABI of libcvl_hip.so."""
import math

import numpy as np


def dtype_of(ty):
    import torch
    return torch.float64 if ty == "d" else torch.int32


def np_dtype(ty):
    return np.float64 if ty == "d" else np.int32


def host(t):
    return t.detach().cpu().numpy()


def eq(a, b, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if a.dtype == np.float64:
        ok = np.array_equal(a.view(np.int64), b.view(np.int64))
    else:
        ok = np.array_equal(a, b)
    if not ok:
        bad = np.nonzero(a != b)[0] if a.dtype != np.float64 else np.nonzero(a.view(np.int64) != b.view(np.int64))[0]
        raise AssertionError(f"{what}: {bad.size} mismatches, first at {bad[:5]}: got {a[bad[:5]]} want {b[bad[:5]]}")


def fp_add_tol(absprefix, seglen):
    """Tolerance for a float +-scan element: 2(ceil(log2 L)+2) 2^-53 sum|x|."""
    L = max(int(seglen), 2)
    return 2.0 * (math.ceil(math.log2(L)) + 2) * 2.0 ** -53 * absprefix


def fp_add_tol_vec(absprefix, seglen):
    """Vector form of fp_add_tol."""
    L = np.maximum(np.asarray(seglen, np.int64), 2)
    return 2.0 * (np.ceil(np.log2(L)) + 2) * 2.0 ** -53 * np.asarray(absprefix)


def seg_lengths_per_element(lens):
    lens = np.asarray(lens, np.int64)
    return np.repeat(lens, lens)
