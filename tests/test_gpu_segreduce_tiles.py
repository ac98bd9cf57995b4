"""Segmented reductions at tile boundaries.

The segmented reductions run in two steps without a look-back: each tile
reduces the segments that start and end in it and records the partial of
the one that enters it from earlier tiles; a second launch completes those
from the tiles' aggregates.  The cases below put segment boundaries exactly
on, just before and just after tile boundaries, make one segment span many
tiles, put empty segments at tile boundaries and pack thousands of
one-element segments around them.  Results must equal the C oracle's
(nesl/cvl/serial/vprims.c:128-201) exactly for integers and min/max, and be
within the compensated-sum bound for double +, through the CVL entry
points and every schedule of cvlhip_seg_reduce (segsched.hpp: flat tiles
of 8, 2 and 1 rows per wave, thread, wave and workgroup per segment), with
the boundaries placed around the tiles of each flat schedule."""
import numpy as np
import pytest

from gpu_util import eq, fp_add_tol_vec, host
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# elements in 512 threads x rows x 16 bytes: a multiple of the tile of a
# segmented reduction in workgroups of 512 threads or fewer
def tile(ty, rows):
    return 8 * rows * 64 * (4 if ty == "z" else 2)


@pytest.fixture(scope="module")
def cv():
    from neslgpu_amd.cvl import Cvl
    return Cvl()


def patterns(T, rng):
    ragged = rng.integers(0, 3 * T, 40)
    return {
        "aligned": [T, T, T, T],
        "spanning": [3, 5 * T + 1, T + 1],
        "near": [T - 1, 2, T - 1, 1, T + 1, T - 2, 3, 2 * T],
        "empties": [T, 0, T, 0, 0, T - 3, 3, 0, T, 0],
        "singletons": [1] * (T + 20) + [T] + [1] * 40,
        "ragged": ragged.tolist(),
    }


def values(ty, n, rng):
    if ty == "d":
        return rng.uniform(-1, 1, n)
    return rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)


@pytest.mark.parametrize("rows", [8, 2, 1])
@pytest.mark.parametrize("ty", ["z", "d"])
def test_segmented_reductions_at_tile_boundaries(cv, ty, rows):
    import torch
    from neslgpu_amd import dispatch as D
    rng = np.random.default_rng((77 if ty == "z" else 78) + rows)
    L = cv.lib
    for name, lens in patterns(tile(ty, rows), rng).items():
        lens = np.asarray(lens, np.int32)
        n, m = int(lens.sum()), lens.size
        x = values(ty, n, rng)
        sd = cv.segd(lens)
        dx = torch.from_numpy(x).to(cv.device)
        out = torch.empty(m, dtype=dx.dtype, device=cv.device)
        for op in ("add", "max", "min") + (("mul",) if ty == "z" else ()):
            if ty == "d" and op == "add":
                want, absum = O.segreduce_add_neumaier(x, lens)
            else:
                want = O.segreduce(op, ty, x, lens)
            runs = [("entry", lambda: getattr(L, f"{op}_re{ty}")(
                out, dx, sd, n, m, cv.scratch(getattr(L, f"{op}_re{ty}_scratch")(n, m))))]
            for arm in range(len(D.ARMS)):
                runs.append((D.ARMS[arm], lambda arm=arm: L.cvlhip_seg_reduce(
                    D.OPS[op], 1 if ty == "d" else 0, arm, out, dx, sd, n, m,
                    cv.scratch(L.cvlhip_seg_scratch(n, m)))))
            for how, run in runs:
                print(name, rows, op, ty, how, n, m, flush=True)
                out.fill_(float("nan") if ty == "d" else 0x5A5A5A5A)   # a segment left unwritten shows
                run()
                got = host(out)
                if ty == "d" and op == "add":
                    assert np.all(np.abs(got - want) <= fp_add_tol_vec(absum, lens)), (name, op, how)
                else:
                    eq(got, want, f"{name} {op}_re{ty} {how}")
    assert L.cvlhip_check_device_errors() == 0


def test_second_step_over_many_rounds(cv):
    """More tiles than the second step folds in one round (8192 records):
    2^27 + 40000 int32 in 8-row tiles of up to 256 threads are more than
    16384 tiles (and the 2- and 1-row tiles many times that), with
    segments of geometric(2^16) lengths crossing tile and round boundaries,
    a few empty ones among them."""
    import torch
    rng = np.random.default_rng(81)
    n = (1 << 27) + 40000
    lens = rng.geometric(1.0 / (1 << 16), n >> 15).astype(np.int64)
    lens[rng.integers(0, lens.size, 40)] = 0
    c = np.cumsum(lens)
    lens = lens[: np.searchsorted(c, n)]
    lens = np.append(lens, n - lens.sum()).astype(np.int32)
    m = lens.size
    x = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
    sd = cv.segd(lens)
    dx = torch.from_numpy(x).to(cv.device)
    out = torch.empty(m, dtype=dx.dtype, device=cv.device)
    L = cv.lib
    from neslgpu_amd import dispatch as D
    for op in ("max", "add"):
        want = O.segreduce(op, "z", x, lens)
        runs = [("entry", lambda: getattr(L, f"{op}_rez")(
            out, dx, sd, n, m, cv.scratch(getattr(L, f"{op}_rez_scratch")(n, m))))]
        for arm in ("flat8", "flat2", "flat1"):     # the tiled schedules
            runs.append((arm, lambda arm=arm: L.cvlhip_seg_reduce(
                D.OPS[op], 0, D.ARMS.index(arm), out, dx, sd, n, m, cv.scratch(L.cvlhip_seg_scratch(n, m)))))
        for how, run in runs:
            print("many rounds", op, how, n, m, flush=True)
            out.fill_(0x5A5A5A5A)
            run()
            eq(host(out), want, f"{op}_rez {how} over many rounds")
    assert L.cvlhip_check_device_errors() == 0
