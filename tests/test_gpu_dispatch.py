"""Segmented-op schedules and the learned dispatcher on the GPU.

Every schedule ("arm") of cvlhip_seg_scan / cvlhip_seg_reduce matches the
oracle: bit-exact for integer and min/max operators, within the compensated
tolerance for double +.  The CVL entries run the arm mke_fov picked for the
descriptor: a network forced to pick each arm in turn leaves that pick in
the descriptor and the entries' results still match the oracle; the pick
the library computes on the device equals the argmin of the HIP forward
pass on the same features.  The network's HIP gradients match the numpy
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
SENTINEL = {"z": 0x5A5A5A5A, "d": float("nan")}     # what an unwritten output element holds


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
                d = torch.full_like(dx, SENTINEL[ty])
                cv.lib.cvlhip_seg_scan(D.OPS[op], 1 if ty == "d" else 0, arm, d, dx, sd, n, m, scr)
                got = host(d)
                if ty == "d" and op == "add":
                    assert np.all(np.abs(got - ref_s) <= tol_s), (D.ARMS[arm], op, ty)
                else:
                    eq(got, want_scan, f"scan {D.ARMS[arm]} {op}{ty} {dist}")
                r = torch.full((m,), SENTINEL[ty], dtype=dx.dtype, device=cv.device)
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


def forcing_params(dims, arm):
    """A network whose prediction is lowest for `arm` whatever the input:
    every weight 0, the output biases 0 except -1 at `arm`."""
    F, H, A, P = dims
    p = np.zeros(P, np.float32)
    p[P - A + arm] = -1.0
    return p


def test_forced_picks_run_through_cvl_entries(cv):
    """Each arm, picked by mke_fov, is what add_sez / add_sed / max_rez /
    min_red / mul_rez run; results match the oracle.  The lengths (1..40,
    with empty segments, 20k segments) admit every arm."""
    import torch
    from neslgpu_amd import dispatch as D
    L = cv.lib
    dims = D.mlp_dims(L)
    saved = torch.empty(dims[3], dtype=torch.float32, device=cv.device)
    L.cvlhip_sched_get_params(saved)
    had = L.cvlhip_sched_has_model()
    rng = np.random.default_rng(21)
    lens = rng.integers(1, 41, size=20_000).astype(np.int32)
    lens = np.insert(lens, [1, 5, 9000], 0)
    n, m = int(lens.sum()), lens.size
    assert n >= L.cvlhip_sched_min_n()
    xi = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
    xd = rng.uniform(-1, 1, n)
    dxi, dxd = torch.from_numpy(xi).to(cv.device), torch.from_numpy(xd).to(cv.device)
    ref_s, absp = O.segscan_add_neumaier(xd, lens)
    tol_s = fp_add_tol_vec(absp, seg_lengths_per_element(lens))
    try:
        for arm in range(len(D.ARMS)):
            L.cvlhip_sched_set_params(torch.from_numpy(forcing_params(dims, arm)).to(cv.device))
            sd = cv.segd(lens)
            for red in (0, 1):
                for eb in (4, 8):
                    assert L.cvlhip_segd_pick(sd, red, eb) == arm, (arm, red, eb)
            print("forced", D.ARMS[arm], n, m, flush=True)
            d = torch.full_like(dxi, SENTINEL["z"])
            L.add_sez(d, dxi, sd, n, m, cv.scratch(L.add_sez_scratch(n, m)))
            eq(host(d), O.segscan("add", "z", xi, lens), f"add_sez {D.ARMS[arm]}")
            d = torch.full_like(dxd, SENTINEL["d"])
            L.add_sed(d, dxd, sd, n, m, cv.scratch(L.add_sed_scratch(n, m)))
            assert np.all(np.abs(host(d) - ref_s) <= tol_s), ("add_sed", D.ARMS[arm])
            for name, x, dx, ty in (("max_rez", xi, dxi, "z"), ("mul_rez", xi, dxi, "z"), ("min_red", xd, dxd, "d")):
                r = torch.full((m,), SENTINEL[ty], dtype=dx.dtype, device=cv.device)
                getattr(L, name)(r, dx, sd, n, m, cv.scratch(getattr(L, name + "_scratch")(n, m)))
                eq(host(r), O.segreduce(name[:3], ty, x, lens), f"{name} {D.ARMS[arm]}")
    finally:
        L.cvlhip_sched_set_params(saved if had else None)
    assert L.cvlhip_check_device_errors() == 0


def test_per_segment_arms_need_short_segments_in_numbers(cv):
    """A network that prefers the thread, wave or workgroup arm gets it only
    where the descriptor admits it (longest segment below the arm's class
    bound, enough segments); elsewhere the pick falls back to flat8."""
    import torch
    from neslgpu_amd import dispatch as D
    L = cv.lib
    dims = D.mlp_dims(L)
    saved = torch.empty(dims[3], dtype=torch.float32, device=cv.device)
    L.cvlhip_sched_get_params(saved)
    had = L.cvlhip_sched_has_model()
    n = 1 << 20
    cases = {                                       # lengths -> admitted per-segment arms
        "two halves": (D.lengths("few", n, 2, None), set()),
        "const 16": (np.full(n // 16, 16, np.int32), {3, 4, 5}),
        "const 100": (np.full(n // 100, 100, np.int32), {4, 5}),
        "const 1000": (np.full(n // 1000, 1000, np.int32), {5}),
        "const 20000": (np.full(n // 20000, 20000, np.int32), set()),
    }
    try:
        for arm in (3, 4, 5):
            L.cvlhip_sched_set_params(torch.from_numpy(forcing_params(dims, arm)).to(cv.device))
            for name, (lens, ok) in cases.items():
                sd = cv.segd(lens)
                got = L.cvlhip_segd_pick(sd, 0, 4)
                assert got == (arm if arm in ok else 0), (name, D.ARMS[arm], got)
    finally:
        L.cvlhip_sched_set_params(saved if had else None)


def test_small_descriptors_carry_no_pick(cv):
    L = cv.lib
    n = L.cvlhip_sched_min_n() - 1
    sd = cv.segd(np.array([n // 2, n - n // 2], np.int32))
    assert L.cvlhip_segd_pick(sd, 0, 4) == -1


def test_device_pick_is_the_networks_argmin(cv):
    """mke_fov's pick equals the argmin of cvlhip_mlp_forward on the
    features cvlhip_segd_features reports, with the library's parameters."""
    import torch
    from neslgpu_amd import dispatch as D
    L = cv.lib
    if not L.cvlhip_sched_has_model():
        pytest.skip("library built without a dispatcher model")
    disp = D.Dispatcher(cv)
    L.cvlhip_sched_get_params(disp.params)
    rng = np.random.default_rng(5)
    for dist, param in (("zipf", 1.5), ("geometric", 4), ("const", 1000), ("uniform", 2000), ("few", 3)):
        for log2n in (15, 18, 21):
            lens = D.lengths(dist, 1 << log2n, param, rng)
            sd = cv.segd(lens)
            for kind, ty in (("scan", "z"), ("scan", "d"), ("reduce", "z"), ("reduce", "d")):
                f = D.features(cv, sd, kind, ty)
                want = disp.choose(f)
                got = L.cvlhip_segd_pick(sd, 1 if kind == "reduce" else 0, 8 if ty == "d" else 4)
                assert got == want, (dist, param, log2n, kind, ty, got, want)


def test_learned_dispatch_gain_on_held_out_workloads(cv):
    """The built-in network on the held-out workloads (dispatch.EVAL_DISTS x
    EVAL_SIZES: Zipf 1.4, 2.0, 2.8 and geometric(512) lengths, 2^17..2^25
    elements; never used in training or in the choice of the training run):
    the CVL entries' total time under the device's picks, plus what the pick
    adds to mke_fov, is within 5% of the exhaustive autotuner's total (the
    fastest arm timed for every case) and below the total of every fixed
    arm."""
    from neslgpu_amd import dispatch as D
    L = cv.lib
    if not L.cvlhip_sched_has_model():
        pytest.skip("library built without a dispatcher model")
    L.cvlhip_sched_enable(1)
    ev = D.evaluate_live(cv, D.EVAL_DISTS, D.EVAL_SIZES, log=lambda s: None)
    print("learned", ev["learned_total_ms"], "decision", ev["decision_total_ms"], "autotuned",
          ev["autotuned_total_ms"], "fixed", ev["fixed_arm_total_ms"], flush=True)
    assert ev["regret"] <= 0.05, ev["regret"]
    for arm, total in ev["fixed_arm_total_ms"].items():
        assert ev["learned_total_ms"] < total, (arm, total, ev["learned_total_ms"])
    assert L.cvlhip_check_device_errors() == 0
