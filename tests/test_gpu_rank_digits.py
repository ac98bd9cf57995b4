"""Radix rank on keys whose digits are partly constant.

A stable LSD pass whose digit takes one value keeps the order, so the rank
skips it (the serial rank, nesl/cvl/serial/rank.c:40-362, runs every
pass; the result is the same).  Between passes the stored keys keep only
the bits the remaining passes read: 24, 16 or 8 bits after digit 0, 1 or
2.  Each case here makes a different set of passes run, so that keys are
narrowed from every position (32 -> 32, 32 -> 16, 32 -> 8, 16 -> 8), and
the ranks must equal the C oracle's exactly, unsegmented and segmented,
up and down, from aligned and unaligned inputs."""
import numpy as np
import pytest

from gpu_util import eq, host
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = 300_007            # many rank tiles and a ragged last one
# above 2^21 keys the inverse of the sort order takes two partition passes,
# and an unsegmented rank's last pass runs the first of them itself
N2 = (1 << 21) + 4101


@pytest.fixture(scope="module")
def cv():
    from neslgpu_amd.cvl import Cvl
    return Cvl()


def keys(kind, n, rng):
    r = rng.integers(-2 ** 31, 2 ** 31, n, dtype=np.int64).astype(np.uint32)
    masks = {
        "all": 0xFFFFFFFF,     # digits 0-3 vary
        "d3": 0xFF000000,      # digit 3 only: one pass, which is also the first
        "d0": 0x000000FF,      # digit 0 only: one pass, pass 0
        "d13": 0xFF00FF00,     # digits 1 and 3 (0 and 2 constant): 32 -> 8
        "d12": 0x00FFFF00,     # digits 1 and 2: 32 -> 16, last pass keyless
        "d03": 0xFF0000FF,     # digits 0 and 3: 32 -> 8
        "d123": 0xFFFFFF00,    # digits 1, 2, 3: 32 -> 16 -> 8
    }
    if kind == "small":        # digits 0 and 1: 32 -> 24 bits kept in 4 bytes
        return rng.integers(0, 1000, n).astype(np.int32)
    return (r & np.uint32(masks[kind])).view(np.int32)


# pass 0 runs first ("all", "d03", "small", "d0") or is skipped (the rest),
# which decides whether its tile counts, taken while the host reads which
# digits vary, are used
KINDS = ["all", "d13", "d12", "d03", "d123", "small", "d3", "d0"]


def device(cv, a, shift):
    import torch
    t = torch.empty(a.size + shift, dtype=torch.int32, device=cv.device)
    v = t[shift:]
    v.copy_(torch.from_numpy(a).to(cv.device))
    return v


@pytest.mark.parametrize("kind", KINDS)
def test_rank_digit_patterns(cv, kind):
    import torch
    rng = np.random.default_rng(KINDS.index(kind) + 100)
    L = cv.lib
    x = keys(kind, N, rng)
    for up in (True, False):
        for shift in (0, 1):                      # 1: the input is not 16-byte aligned
            name = f"rk{'u' if up else 'd'}_luz"
            print(name, kind, "shift", shift, flush=True)
            d = torch.empty(N, dtype=torch.int32, device=cv.device)
            getattr(L, name)(d, device(cv, x, shift), N, cv.scratch(getattr(L, name + "_scratch")(N)))
            eq(host(d), O.rank(x, up), f"{name} {kind} shift {shift}")
    assert L.cvlhip_check_device_errors() == 0


@pytest.mark.parametrize("kind", ["d13", "d123", "small"])
def test_segmented_rank_digit_patterns(cv, kind):
    import torch
    rng = np.random.default_rng(KINDS.index(kind) + 200)
    L = cv.lib
    x = keys(kind, N, rng)
    cuts = np.sort(rng.integers(0, N + 1, 700))
    lens = np.diff(np.concatenate([[0], cuts, [N]])).astype(np.int32)
    sd = cv.segd(lens)
    for up in (True, False):
        name = f"rk{'u' if up else 'd'}_lez"
        print(name, kind, flush=True)
        d = torch.empty(N, dtype=torch.int32, device=cv.device)
        getattr(L, name)(d, device(cv, x, 0), sd, N, lens.size,
                         cv.scratch(getattr(L, name + "_scratch")(N, lens.size)))
        eq(host(d), O.rank(x, up, lens), f"{name} {kind}")
    assert L.cvlhip_check_device_errors() == 0


FUSED_KINDS = ["all", "d12", "d3", "small", "d0"]


@pytest.mark.parametrize("kind", FUSED_KINDS)
def test_rank_last_pass_fused_with_inverse(cv, kind):
    """Above 2^21 keys: the last pass places (index, position) pairs for the
    inverse's first partition pass instead of writing the order -- with
    every pass running, with the last pass also the first (digit 3 or digit
    0 only: indices are positions; for digit 0 the pass also counts the
    inverse's children, so pass 0's early counts go unused), with a keyless
    last pass, and for doubles with ties, up and down, aligned and not."""
    import torch
    rng = np.random.default_rng(FUSED_KINDS.index(kind) + 300)
    L = cv.lib
    x = keys(kind, N2, rng)
    for up in (True, False):
        for shift in (0, 1):
            name = f"rk{'u' if up else 'd'}_luz"
            print(name, kind, N2, "shift", shift, flush=True)
            d = torch.empty(N2, dtype=torch.int32, device=cv.device)
            getattr(L, name)(d, device(cv, x, shift), N2, cv.scratch(getattr(L, name + "_scratch")(N2)))
            eq(host(d), O.rank(x, up), f"{name} {kind} n {N2} shift {shift}")
    if kind == "all":
        y = np.round(rng.uniform(-1e3, 1e3, N2), 1)          # many ties
        dy = torch.from_numpy(y).to(cv.device)
        for up in (True, False):
            name = f"rk{'u' if up else 'd'}_lud"
            print(name, N2, flush=True)
            d = torch.empty(N2, dtype=torch.int32, device=cv.device)
            getattr(L, name)(d, dy, N2, cv.scratch(getattr(L, name + "_scratch")(N2)))
            eq(host(d), O.rank(y, up), f"{name} n {N2}")
    assert L.cvlhip_check_device_errors() == 0
