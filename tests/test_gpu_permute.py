"""Permute parity across the sizes where the partitioned permute changes
shape: a single LDS block (n <= 8192), one partition pass (n <= 2^21) and
two passes, for int, bool and double, plain and segmented; plus the
out-of-range guard (indices outside the vector are dropped and reported,
never written)."""
import numpy as np
import pytest

from gpu_util import eq, host
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIZES = [1, 7, 8191, 8192, 8193, 100003, 1 << 21, (1 << 21) + 5, 3_000_017]


@pytest.fixture(scope="module")
def cv():
    from neslgpu_amd.cvl import Cvl
    return Cvl()


def dev(cv, a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(cv.device)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("ty", ["z", "d"])
def test_smp_pu_sizes(cv, n, ty):
    import torch
    rng = np.random.default_rng(n)
    p = rng.permutation(n).astype(np.int32)
    s = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32) if ty == "z" else rng.uniform(-1, 1, n)
    print("smp_pu", ty, n, flush=True)
    d = torch.zeros(n, dtype=torch.int32 if ty == "z" else torch.float64, device=cv.device)
    fn = getattr(cv.lib, f"smp_pu{ty}")
    fn(d, dev(cv, s), dev(cv, p), n, cv.scratch(getattr(cv.lib, f"smp_pu{ty}_scratch")(n)))
    eq(host(d), O.smp_pu(ty, s, p), f"smp_pu{ty} {n}")
    assert cv.lib.cvlhip_check_device_errors() == 0


@pytest.mark.parametrize("n", [5000, 300_000, 5_000_000])
def test_smp_pe_sizes(cv, n):
    import torch
    rng = np.random.default_rng(n + 1)
    lens = rng.geometric(1 / 300, size=n // 300 + 100).astype(np.int64)
    lens = lens[np.cumsum(lens) <= n]
    lens = np.append(lens, n - lens.sum()).astype(np.int32)
    lens = np.insert(lens, 3, 0)          # an empty segment too
    idx = np.concatenate([rng.permutation(int(L)) for L in lens]).astype(np.int32)
    s = rng.integers(0, 1 << 30, n).astype(np.int32)
    print("smp_pe", n, lens.size, flush=True)
    sd = cv.segd(lens)
    d = torch.zeros(n, dtype=torch.int32, device=cv.device)
    cv.lib.smp_pez(d, dev(cv, s), dev(cv, idx), sd, n, lens.size,
                   cv.scratch(cv.lib.smp_pez_scratch(n, lens.size)))
    eq(host(d), O.smp_pe("z", s, idx, lens), f"smp_pez {n}")
    assert cv.lib.cvlhip_check_device_errors() == 0


@pytest.mark.parametrize("n", [1000, 200_000])
def test_out_of_range_index_is_dropped(cv, n):
    """Out-of-range indices are skipped and reported; nothing outside d is
    written (a guard region after d keeps its sentinel)."""
    import torch
    rng = np.random.default_rng(3)
    p = rng.permutation(n).astype(np.int32)
    p[5] = n + 1000
    p[7] = -3
    s = np.arange(n, dtype=np.int32)
    buf = torch.full((n + 4096,), -7, dtype=torch.int32, device=cv.device)
    d = buf[:n]
    print("out-of-range", n, flush=True)
    cv.lib.smp_puz(d, dev(cv, s), dev(cv, p), n, cv.scratch(cv.lib.smp_puz_scratch(n)))
    assert cv.lib.cvlhip_check_device_errors() != 0
    assert np.all(host(buf[n:]) == -7)
    got = host(d)
    ok = (p >= 0) & (p < n)
    assert np.array_equal(got[p[ok]], s[ok])
