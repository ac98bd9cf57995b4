"""cvlhip_add_sez_noalloc: add_sez with its scratch in a static device array,
for callers that cannot allocate.  Its results must equal add_sez's bit for
bit on two calls in a row, the second on a smaller input (the array keeps
the first call's contents), through both schedules it can take (the fixed
one below 2^15 elements, the dispatched one above); n beyond the array's
capacity is refused before any work is queued."""
import numpy as np
import pytest

from gpu_util import eq, host
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NOALLOC_MAX_N = 1073545216


@pytest.fixture(scope="module")
def cv():
    from neslgpu_amd.cvl import Cvl
    return Cvl()


def run_both(cv, lens, rng):
    import torch
    L = cv.lib
    n, m = int(lens.sum()), lens.size
    x = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
    dx = torch.from_numpy(x).to(cv.device)
    sd = cv.segd(lens)
    ref = torch.full_like(dx, 0x5A5A5A5A)
    L.add_sez(ref, dx, sd, n, m, cv.scratch(L.add_sez_scratch(n, m)))
    got = torch.full_like(dx, 0x3C3C3C3C)
    L.cvlhip_add_sez_noalloc(got, dx, sd, n, m)
    eq(host(got), host(ref), f"noalloc vs add_sez n={n} m={m}")
    eq(host(got), O.segscan("add", "z", x, lens), f"noalloc vs oracle n={n} m={m}")


def test_noalloc_matches_add_sez_twice_second_smaller(cv):
    from neslgpu_amd import dispatch as D
    rng = np.random.default_rng(31)
    big = D.lengths("zipf", 3_000_017, 1.5, rng, lmax=1 << 16)
    small = D.lengths("geometric", 20_011, 40, rng)
    for lens in (big, small):
        print("noalloc", int(lens.sum()), lens.size, flush=True)
        run_both(cv, lens, rng)
    one = np.array([50_000], np.int32)                 # one segment: the unsegmented scan
    run_both(cv, one, rng)
    assert cv.lib.cvlhip_check_device_errors() == 0


def test_noalloc_refuses_n_beyond_its_array(cv):
    from neslgpu_amd.cvl import CvlError
    with pytest.raises(CvlError, match="more than its"):
        cv.lib.cvlhip_add_sez_noalloc(None, None, None, NOALLOC_MAX_N + 1, 2)
