"""The primitives on vectors that are not 16-byte aligned.

CVL vectors are handles into the caller's arena at 8-byte-unit offsets
(nesl/vcode/vstack.c), so a vector may start anywhere a unit starts, and
int vectors can be 4-byte aligned when a caller packs them.  The kernels
read and write 16 bytes per lane only when the pointers allow it and fall
back to narrower accesses otherwise; here every input and output starts
one element past an aligned address, and results must equal the oracle's
exactly (double sums: within the compensated-sum bound)."""
import numpy as np
import pytest

from gpu_util import eq, fp_add_tol_vec, host, seg_lengths_per_element
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = 300_007          # several tiles of every kernel, and a ragged end


@pytest.fixture(scope="module")
def cv():
    from neslgpu_amd.cvl import Cvl
    return Cvl()


def shifted(cv, a):
    """A device copy of `a` starting one element past a 256-byte boundary."""
    import torch
    a = np.ascontiguousarray(a)
    t = torch.empty(a.size + 1, dtype=torch.from_numpy(a[:0]).dtype, device=cv.device)
    v = t[1:]
    v.copy_(torch.from_numpy(a).to(cv.device))
    assert v.data_ptr() % 16 != 0
    return v


def empty_shifted(cv, n, dtype):
    import torch
    return torch.empty(n + 1, dtype=dtype, device=cv.device)[1:]


def lens_for(n, rng):
    cuts = np.sort(rng.integers(0, n + 1, n // 200))
    return np.diff(np.concatenate([[0], cuts, [n]])).astype(np.int32)


def test_scans_and_reductions(cv):
    import torch
    rng = np.random.default_rng(41)
    L = cv.lib
    xi = rng.integers(-2 ** 31, 2 ** 31, N).astype(np.int32)
    xd = rng.uniform(-1, 1, N)
    lens = lens_for(N, rng)
    sd = cv.segd(lens)
    m = lens.size
    scr = cv.scratch(max(L.add_sez_scratch(N, m), L.add_sed_scratch(N, m), L.max_ruz_scratch(N)))
    for ty, x in (("z", xi), ("d", xd)):
        dx = shifted(cv, x)
        tdt = torch.float64 if ty == "d" else torch.int32
        for op in ("add", "max"):
            print("unaligned", op, ty, flush=True)
            d = empty_shifted(cv, N, tdt)
            getattr(L, f"{op}_su{ty}")(d, dx, N, scr)
            if ty == "d" and op == "add":
                ref, absp = O.segscan_add_neumaier(x, np.array([N], np.int32))
                assert np.all(np.abs(host(d) - ref) <= fp_add_tol_vec(absp, np.full(N, N)))
            else:
                eq(host(d), O.scan(op, ty, x), f"{op}_su{ty}")
            getattr(L, f"{op}_se{ty}")(d, dx, sd, N, m, scr)
            if ty == "d" and op == "add":
                ref, absp = O.segscan_add_neumaier(x, lens)
                assert np.all(np.abs(host(d) - ref) <= fp_add_tol_vec(absp, seg_lengths_per_element(lens)))
            else:
                eq(host(d), O.segscan(op, ty, x, lens), f"{op}_se{ty}")
            r = getattr(L, f"{op}_ru{ty}")(dx, N, scr)
            want = O.reduce(op, ty, x)
            if ty == "d" and op == "add":
                assert abs(r - float(want)) <= 1e-9 * np.abs(x).sum()
            else:
                assert r == want, (op, ty)


def test_pack(cv):
    import torch
    rng = np.random.default_rng(42)
    L = cv.lib
    for ty in ("z", "d"):
        x = rng.integers(-2 ** 31, 2 ** 31, N).astype(np.int32) if ty == "z" else rng.uniform(-1, 1, N)
        f = (rng.random(N) < 0.4).astype(np.int32)
        kept = int(f.sum())
        d = empty_shifted(cv, kept, torch.float64 if ty == "d" else torch.int32)
        getattr(L, f"pk2_lu{ty}")(d, shifted(cv, x), shifted(cv, f), N, kept,
                                  cv.scratch(getattr(L, f"pk2_lu{ty}_scratch")(N)))
        eq(host(d), O.pk2(ty, x, f), f"pk2_lu{ty}")


def test_permute_and_back_permute(cv):
    import torch
    rng = np.random.default_rng(43)
    L = cv.lib
    for ty in ("z", "d"):
        x = rng.integers(-2 ** 31, 2 ** 31, N).astype(np.int32) if ty == "z" else rng.uniform(-1, 1, N)
        p = rng.permutation(N).astype(np.int32)
        tdt = torch.float64 if ty == "d" else torch.int32
        d = empty_shifted(cv, N, tdt)
        getattr(L, f"smp_pu{ty}")(d, shifted(cv, x), shifted(cv, p), N,
                                  cv.scratch(getattr(L, f"smp_pu{ty}_scratch")(N)))
        eq(host(d), O.smp_pu(ty, x, p), f"smp_pu{ty}")
        b = empty_shifted(cv, N, tdt)
        getattr(L, f"bck_pu{ty}")(b, shifted(cv, x), shifted(cv, p), N, N,
                                  cv.scratch(getattr(L, f"bck_pu{ty}_scratch")(N, N)))
        eq(host(b), O.bck_pu(ty, x, p), f"bck_pu{ty}")
    assert L.cvlhip_check_device_errors() == 0


def test_rank(cv):
    import torch
    rng = np.random.default_rng(44)
    L = cv.lib
    for ty in ("z", "d"):
        x = rng.integers(-2 ** 31, 2 ** 31, N).astype(np.int32) if ty == "z" else rng.uniform(-1, 1, N)
        for up in (True, False):
            d = empty_shifted(cv, N, torch.int32)
            name = f"rk{'u' if up else 'd'}_lu{ty}"
            getattr(L, name)(d, shifted(cv, x), N, cv.scratch(getattr(L, name + "_scratch")(N)))
            eq(host(d), O.rank(x, up), name)
