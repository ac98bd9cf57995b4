"""Scans and pack at sizes around their tile sizes.

Tile shapes differ by case: plain 4-byte scans take 1024 threads x 12 rows
(49152 elements), segmented 4-byte scans 512 x 8 (16384), double scans
1024 x 8 (16384), pack 1024 x 8 (32768 ints, 16384 doubles).  Each size
here is one tile, one element either side of it, and several tiles with a
ragged end, so the last tile's bounds checks and the look-back over a few
tiles both run.  Results must equal the C oracle's exactly (double sums:
within the compensated-sum bound)."""
import numpy as np
import pytest

from gpu_util import eq, fp_add_tol_vec, host, seg_lengths_per_element
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TILES = {("z", False): 49152, ("z", True): 16384, ("d", False): 16384, ("d", True): 16384}
PACK_TILES = {"z": 32768, "d": 16384}


@pytest.fixture(scope="module")
def cv():
    from neslgpu_amd.cvl import Cvl
    return Cvl()


def sizes(T):
    return [T - 1, T, T + 1, 3 * T + 17]


def vals(ty, n, rng):
    return rng.uniform(-1, 1, n) if ty == "d" else rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)


@pytest.mark.parametrize("ty", ["z", "d"])
def test_scans_around_tiles(cv, ty):
    import torch
    rng = np.random.default_rng(90 if ty == "z" else 91)
    L = cv.lib
    for seg in (False, True):
        for n in sizes(TILES[(ty, seg)]):
            x = vals(ty, n, rng)
            dx = torch.from_numpy(x).to(cv.device)
            d = torch.empty_like(dx)
            if seg:
                cuts = np.sort(rng.integers(0, n + 1, max(1, n // 300)))
                lens = np.diff(np.concatenate([[0], cuts, [n]])).astype(np.int32)
            else:
                lens = np.array([n], np.int32)
            sd = cv.segd(lens)
            m = lens.size
            for op in ("add", "max"):
                name = f"{op}_s{'e' if seg else 'u'}{ty}"
                print(name, n, flush=True)
                if seg:
                    getattr(L, name)(d, dx, sd, n, m, cv.scratch(getattr(L, name + "_scratch")(n, m)))
                else:
                    getattr(L, name)(d, dx, n, cv.scratch(getattr(L, name + "_scratch")(n)))
                got = host(d)
                if ty == "d" and op == "add":
                    ref, absp = O.segscan_add_neumaier(x, lens)
                    assert np.all(np.abs(got - ref) <= fp_add_tol_vec(absp, seg_lengths_per_element(lens))), (name, n)
                else:
                    want = O.segscan(op, ty, x, lens) if seg else O.scan(op, ty, x)
                    eq(got, want, f"{name} n={n}")
    assert L.cvlhip_check_device_errors() == 0


@pytest.mark.parametrize("ty", ["z", "d"])
def test_pack_around_tiles(cv, ty):
    import torch
    rng = np.random.default_rng(92 if ty == "z" else 93)
    L = cv.lib
    for n in sizes(PACK_TILES[ty]):
        x = vals(ty, n, rng)
        f = (rng.random(n) < 0.5).astype(np.int32)
        kept = int(f.sum())
        d = torch.empty(kept, dtype=torch.float64 if ty == "d" else torch.int32, device=cv.device)
        print("pk2", ty, n, flush=True)
        getattr(L, f"pk2_lu{ty}")(d, torch.from_numpy(x).to(cv.device), torch.from_numpy(f).to(cv.device), n, kept,
                                  cv.scratch(getattr(L, f"pk2_lu{ty}_scratch")(n)))
        eq(host(d), O.pk2(ty, x, f), f"pk2_lu{ty} n={n}")
    assert L.cvlhip_check_device_errors() == 0
