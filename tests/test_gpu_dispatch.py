"""Segmented-op schedules and the learned dispatcher on the GPU.

Every schedule ("arm") of cvlhip_seg_scan / cvlhip_seg_reduce matches the
oracle: bit-exact for integer and min/max operators, within the compensated
tolerance for double +.  The network's HIP gradients match the numpy
restatement (itself pinned by finite differences), and training on timings
lowers the loss."""
import numpy as np
import pytest

import mlp_ref
from gpu_util import eq, fp_add_tol_vec, host, seg_lengths_per_element
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cv():
    from neslgpu_amd.cvl import Cvl
    return Cvl()


DISTS = [("geometric", 3), ("geometric", 300), ("zipf", 1.5), ("const", 1), ("bimodal", 0)]


@pytest.mark.parametrize("dist", DISTS)
def test_every_arm_matches_oracle(cv, dist):
    import torch
    from neslgpu_amd import dispatch as D
    rng = np.random.default_rng(7)
    n = 300_000
    lens = D.lengths(dist[0], n, dist[1], rng, lmax=1 << 17)
    lens = np.insert(lens, 2, 0)                        # an empty segment
    m = lens.size
    sd = cv.segd(lens)
    xi = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
    xd = rng.uniform(-1, 1, n)
    scr = cv.scratch(cv.lib.cvlhip_seg_scratch(n, m))
    ref_s, absp = O.segscan_add_neumaier(xd, lens)
    tol_s = fp_add_tol_vec(absp, seg_lengths_per_element(lens))
    ref_r, absum = O.segreduce_add_neumaier(xd, lens)
    tol_r = fp_add_tol_vec(absum, lens)
    for ty, x in (("z", xi), ("d", xd)):
        dx = torch.from_numpy(x).to(cv.device)
        for op in ("add", "min", "max"):
            want_scan = O.segscan(op, ty, x, lens)
            want_red = O.segreduce(op, ty, x, lens)
            for arm in range(len(D.ARMS)):
                print("arm", D.ARMS[arm], dist, ty, op, flush=True)
                d = torch.empty_like(dx)
                cv.lib.cvlhip_seg_scan(D.OPS[op], 1 if ty == "d" else 0, arm, d, dx, sd, n, m, scr)
                got = host(d)
                if ty == "d" and op == "add":
                    assert np.all(np.abs(got - ref_s) <= tol_s), (D.ARMS[arm], op, ty)
                else:
                    eq(got, want_scan, f"scan {D.ARMS[arm]} {op}{ty} {dist}")
                r = torch.empty(m, dtype=dx.dtype, device=cv.device)
                cv.lib.cvlhip_seg_reduce(D.OPS[op], 1 if ty == "d" else 0, arm, r, dx, sd, n, m, scr)
                got = host(r)
                if ty == "d" and op == "add":
                    assert np.all(np.abs(got - ref_r) <= tol_r), (D.ARMS[arm], op, ty)
                else:
                    eq(got, want_red, f"reduce {D.ARMS[arm]} {op}{ty} {dist}")


def test_double_add_scan_is_deterministic(cv):
    """The flat schedule's double +-scan gives the same bits on every run."""
    import torch
    n = 1 << 22
    x = torch.empty(n, dtype=torch.float64, device=cv.device)
    cv.lib.cvlhip_fill_splitmix_f64(x, n, 99, -1.0, 1.0)
    scr = cv.scratch(cv.lib.add_sud_scratch(n))
    outs = []
    for _ in range(4):
        d = torch.empty_like(x)
        cv.lib.add_sud(d, x, n, scr)
        outs.append(host(d).view(np.int64))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_mlp_gradients_match_numpy(cv):
    import torch
    from neslgpu_amd import dispatch as D
    disp = D.Dispatcher(cv, seed=3)
    F, H, A = disp.F, disp.H, disp.A
    rng = np.random.default_rng(1)
    B = 37
    X = rng.random((B, F)).astype(np.float32)
    T = rng.normal(0, 1, (B, A)).astype(np.float32)
    M = (rng.random((B, A)) < 0.9).astype(np.float32)
    loss, g = disp.gradients(X, T, M)
    p = disp.params.cpu().numpy().astype(np.float64)
    ref_loss, ref_g = mlp_ref.loss_and_grad(p, X.astype(np.float64), T.astype(np.float64),
                                            M.astype(np.float64), F, H, A)
    assert abs(loss - ref_loss) <= 1e-4 * max(1.0, abs(ref_loss))
    gg = g.cpu().numpy()
    assert np.allclose(gg, ref_g, rtol=2e-3, atol=2e-5), float(np.abs(gg - ref_g).max())
    pred = disp.predict(X)
    ref_pred, _ = mlp_ref.forward(p, X.astype(np.float64), F, H, A)
    assert np.allclose(pred, ref_pred, rtol=1e-4, atol=1e-5)


def test_training_lowers_loss(cv):
    from neslgpu_amd import dispatch as D
    disp = D.Dispatcher(cv, seed=0)
    rng = np.random.default_rng(2)
    X = rng.random((128, disp.F)).astype(np.float32)
    W = rng.normal(0, 1, (disp.F, disp.A)).astype(np.float32)
    T = (X @ W).astype(np.float32)
    M = np.ones_like(T)
    losses = disp.train(X, T, M, epochs=600, lr=3e-3, batch=64)
    assert np.mean(losses[-20:]) < 0.2 * np.mean(losses[:20])
