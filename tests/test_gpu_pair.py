"""add_sed and max_red of the same values in one pass
(cvlhip_add_sed_max_red), the irregular float64 workload's two operations.

The reduction must equal max_red bit for bit (and the oracle's fold); the
scan must lie within the compensated-sum bound of the oracle's add_sed, as
the separate add_sed does, and give the same bits on every run.  Segment
shapes: Zipf lengths with empty segments, one segment, and no elements."""
import numpy as np
import pytest

from gpu_util import eq, fp_add_tol_vec, host, seg_lengths_per_element
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cv():
    from neslgpu_amd.cvl import Cvl
    return Cvl()


def shapes():
    from neslgpu_amd import dispatch as D
    rng = np.random.default_rng(11)
    z = D.lengths("zipf", 1 << 20, 1.5, rng, lmax=1 << 16)
    z = np.insert(z, [0, 3, z.size], 0)                     # empty segments, first and last too
    g = D.lengths("geometric", 100_003, 40, rng, lmax=1 << 12)
    return [("zipf", z), ("geometric", g), ("one", np.array([70_001], np.int32)),
            ("tiny", np.array([0, 1, 0, 2], np.int32)), ("empty", np.zeros(5, np.int32))]


@pytest.mark.parametrize("name,lens", shapes(), ids=[s[0] for s in shapes()])
def test_pair_matches_separate_ops(cv, name, lens):
    import torch
    lens = np.asarray(lens, np.int32)
    n, m = int(lens.sum()), lens.size
    rng = np.random.default_rng(m)
    x = rng.uniform(-1, 1, n)
    print("pair", name, n, m, flush=True)
    sd = cv.segd(lens)
    dx = torch.from_numpy(x).to(cv.device)
    scr = cv.scratch(max(cv.lib.cvlhip_add_sed_max_red_scratch(n, m), cv.lib.max_red_scratch(n, m)))
    runs = []
    for _ in range(2):
        scan = torch.empty(n, dtype=torch.float64, device=cv.device)
        red = torch.empty(m, dtype=torch.float64, device=cv.device)
        cv.lib.cvlhip_add_sed_max_red(scan, red, dx, sd, n, m, scr)
        runs.append((host(scan), host(red)))
    scan, red = runs[0]
    assert np.array_equal(scan.view(np.int64), runs[1][0].view(np.int64)), "scan bits differ between runs"
    want_red = torch.empty(m, dtype=torch.float64, device=cv.device)
    cv.lib.max_red(want_red, dx, sd, n, m, scr)
    eq(red, host(want_red), f"pair max vs max_red {name}")
    eq(red, O.segreduce("max", "d", x, lens), f"pair max vs oracle {name}")
    if n:
        ref, absp = O.segscan_add_neumaier(x, lens)
        tol = fp_add_tol_vec(absp, seg_lengths_per_element(lens))
        assert np.all(np.abs(scan - ref) <= tol), name
        # heads of segments start from the identity exactly
        offs = np.concatenate([[0], np.cumsum(lens)[:-1]])[lens > 0]
        assert np.all(scan[offs] == 0.0)
    assert cv.lib.cvlhip_check_device_errors() == 0
