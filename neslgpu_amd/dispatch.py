"""Scheduler for segmented scans and reductions: exhaustive autotuner and a
learned dispatcher.

A segmented operation can run under several schedules ("arms",
csrc/segsched.hpp): flat look-back tiles of 8, 2 or 1 rows per wave, or
thread-, wave- or workgroup-per-segment.  The autotuner times every arm on a
workload (HIP events).  The dispatcher is a small MLP that reads the
descriptor statistics mke_fov computes on the device (n, m, log2-length
histograms by elements and by segments, value size, scan/reduce) and
predicts every arm's log-time.  The library runs the network itself, in
mke_fov, with the parameters trained here (built into libcvl_hip.so from
dispatch_model.npz), and stores the pick in the descriptor; every CVL
segmented call on that descriptor runs the picked arm, with no host round
trip.  The network's forward and backward passes and its Adam steps are HIP
kernels (cvlhip_mlp_*); this module only orchestrates them.  Only the
schedule changes with the choice, never the semantics; the autotuner's
exhaustive timings stay the measure of the dispatcher's regret.

Three workload sets, kept apart:
  TRAIN  -- timings the network is fitted to;
  VALID  -- picks the training run (epochs, seed) with the lowest regret;
  EVAL   -- held out: never used for any choice; its regret is the figure
            reported.

    python -m neslgpu_amd.dispatch train [record.json] [model.npz]
        # gather timings, train, save the model, evaluate on EVAL live
    python -m neslgpu_amd.dispatch eval [record.json]
        # the library's built-in model on EVAL, live
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "dispatch_model.npz")

ARMS = ["flat8", "flat2", "flat1", "thread", "wave", "block"]
OPS = {"add": 0, "mul": 1, "min": 2, "max": 3}
# (kind, operator, value type): one of each kind of pick a descriptor holds
KINDS = (("scan", "add", "z"), ("scan", "add", "d"), ("reduce", "max", "z"), ("reduce", "max", "d"))
CVL_NAME = {("scan", "add", "z"): "add_sez", ("scan", "add", "d"): "add_sed",
            ("reduce", "max", "z"): "max_rez", ("reduce", "max", "d"): "max_red"}


def mlp_dims(lib):
    a = [C.c_int() for _ in range(4)]
    lib.dll.cvlhip_mlp_dims(*[C.byref(x) for x in a])
    return tuple(x.value for x in a)


# ------------------------------------------------------------------ workloads
def lengths(kind: str, n: int, param: float, rng: np.random.Generator, lmax: int = 1 << 20) -> np.ndarray:
    """Segment lengths summing exactly to n ("few": n split into `param`
    equal segments, not clipped to lmax)."""
    if kind == "few":
        k = int(param)
        L = np.full(k, n // k, np.int64)
        L[: n % k] += 1
        return L[L > 0].astype(np.int32)
    out, tot = [], 0
    while tot < n:
        k = max(1024, int(2 * (n - tot) / max(param, 1.0)) + 16)
        if kind == "geometric":
            z = rng.geometric(1.0 / param, size=k)
        elif kind == "uniform":
            z = rng.integers(1, int(2 * param) + 1, size=k)
        elif kind == "zipf":
            z = rng.zipf(param, size=k)
        elif kind == "const":
            z = np.full(k, int(param))
        elif kind == "bimodal":            # many short segments and a few long ones
            z = np.where(rng.random(k) < 0.002, rng.integers(1 << 14, 1 << 18, size=k), rng.integers(1, 8, size=k))
        else:
            raise ValueError(kind)
        z = np.minimum(z, lmax)
        out.append(z)
        tot += int(z.sum())
    L = np.concatenate(out).astype(np.int64)
    cs = np.cumsum(L)
    k = int(np.searchsorted(cs, n))
    L = L[: k + 1].copy()
    L[-1] -= int(cs[k] - n)
    return L[L > 0].astype(np.int32)


TRAIN_DISTS = ([("geometric", m) for m in (2, 8, 64, 1024, 8192)]
               + [("uniform", m) for m in (3, 200, 3000)]
               + [("zipf", a) for a in (1.3, 1.7, 2.2, 3.0)]
               + [("const", c) for c in (1, 32, 100000)]
               + [("bimodal", 0)]
               + [("few", k) for k in (2, 16, 300)])
TRAIN_SIZES = [1 << k for k in (15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25)]
VALID_DISTS = [("zipf", 1.6), ("geometric", 256), ("uniform", 40)]
VALID_SIZES = [1 << k for k in (16, 18, 20, 22, 24)]
EVAL_DISTS = [("zipf", 1.4), ("zipf", 2.0), ("zipf", 2.8), ("geometric", 512)]
EVAL_SIZES = [1 << k for k in (17, 19, 21, 23, 25)]


# ------------------------------------------------------------------ timing
class Timer:
    """Median time (ms) of a callable on the library's stream, HIP events."""

    def __init__(self, reps: int = 5, slow_ms: float = 50.0):
        import torch
        self.torch, self.reps, self.slow_ms = torch, reps, slow_ms

    def __call__(self, fn) -> float:
        torch = self.torch
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        times = []
        for r in range(self.reps + 1):
            torch.cuda.synchronize()
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            t = e0.elapsed_time(e1)
            if r > 0:
                times.append(t)
            if t > self.slow_ms:          # a hopeless arm: one timing is enough
                return t
        return float(np.median(times))


class Autotuner:
    """Times every arm of a segmented op."""

    def __init__(self, cv, reps: int = 5):
        self.cv, self.timer = cv, Timer(reps)

    def run(self, kind, op, ty, arm, d, s, sd, scr):
        L = self.cv.lib
        f = L.cvlhip_seg_scan if kind == "scan" else L.cvlhip_seg_reduce
        f(OPS[op], 1 if ty == "d" else 0, arm, d, s, sd, sd.n, sd.m, scr)

    def tune(self, kind, op, ty, d, s, sd) -> np.ndarray:
        scr = self.cv.scratch(self.cv.lib.cvlhip_seg_scratch(sd.n, sd.m))
        return np.array([self.timer(lambda a=a: self.run(kind, op, ty, a, d, s, sd, scr))
                         for a in range(len(ARMS))])


def features(cv, sd, kind: str, ty: str) -> np.ndarray:
    """The network's input for this descriptor and kind of operation, as the
    library computes it (from the histogram mke_fov stored)."""
    F = cv.lib.cvlhip_seg_features_count()
    out = np.zeros(F, np.float32)
    cv.lib.cvlhip_segd_features(sd, sd.n, sd.m, out)
    out[F - 2] = 1.0 if ty == "d" else 0.0
    out[F - 1] = 1.0 if kind == "reduce" else 0.0
    return out


# Schedules a descriptor admits (csrc/mlp.hpp sched_admitted): the thread,
# wave and workgroup arms only below a longest-segment class and above a
# segment count; the flat arms always.
ARM_MAX_CLASS = {3: 5, 4: 12, 5: 15}
ARM_MIN_LOG_M = {3: 14, 4: 11, 5: 9}


def admitted(feat: np.ndarray) -> np.ndarray:
    """Boolean mask over ARMS for feature vectors (rows of `feat`), from
    the segment-class fractions and log2(m)/32 the library computes."""
    feat = np.atleast_2d(feat)
    classes = (feat.shape[1] - 5) // 2
    segs = feat[:, 3 + classes: 3 + 2 * classes]
    present = segs > 0
    cmax = np.where(present.any(axis=1), classes - 1 - np.argmax(present[:, ::-1], axis=1), -1)
    logm = feat[:, 1].astype(np.float64) * 32.0
    ok = np.ones((feat.shape[0], len(ARMS)), bool)
    for a in ARM_MAX_CLASS:
        ok[:, a] = (cmax <= ARM_MAX_CLASS[a]) & (logm >= ARM_MIN_LOG_M[a] - 1e-6)
    return ok


def masked_argmin(pred: np.ndarray, ok: np.ndarray) -> np.ndarray:
    """First index of the smallest admitted prediction per row."""
    return np.where(ok, pred, np.inf).argmin(axis=1)


# ------------------------------------------------------------------ the network
class Dispatcher:
    """MLP over descriptor statistics -> predicted log-time per arm."""

    def __init__(self, cv, seed: int = 0):
        import torch
        self.cv, self.torch = cv, torch
        self.F, self.H, self.A, self.P = mlp_dims(cv.lib)
        g = torch.Generator().manual_seed(seed)
        p = torch.zeros(self.P, dtype=torch.float32)
        o = 0
        for fan_in, fan_out in ((self.F, self.H), (self.H, self.H), (self.H, self.A)):
            w = torch.randn(fan_out * fan_in, generator=g) * math.sqrt(2.0 / fan_in)
            p[o:o + w.numel()] = w
            o += w.numel() + fan_out          # biases start at 0
        self.params = p.to(cv.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0

    def _dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(self.cv.device)

    def predict(self, X: np.ndarray) -> np.ndarray:
        torch = self.torch
        B = X.shape[0]
        Xd = self._dev(X)
        out = torch.empty(B * self.A, dtype=torch.float32, device=self.cv.device)
        acts = torch.empty(2 * B * self.H, dtype=torch.float32, device=self.cv.device)
        self.cv.lib.cvlhip_mlp_forward(self.params, Xd, B, out, acts)
        return out.cpu().numpy().reshape(B, self.A)

    def gradients(self, X, T, M):
        """(loss, grads) of the masked MSE on one batch, from the HIP kernels."""
        torch = self.torch
        B = X.shape[0]
        grads = torch.zeros(self.P, dtype=torch.float32, device=self.cv.device)
        loss = torch.zeros(1, dtype=torch.float32, device=self.cv.device)
        acts = torch.empty(2 * B * self.H, dtype=torch.float32, device=self.cv.device)
        work = torch.empty(B * (2 * self.H + 2 * self.A), dtype=torch.float32, device=self.cv.device)
        self.cv.lib.cvlhip_mlp_gradients(self.params, self._dev(X), self._dev(T), self._dev(M), B,
                                         float(np.sum(M)), grads, loss, acts, work)
        return float(loss.item()), grads

    def train(self, X, T, M, epochs: int = 2000, lr: float = 3e-3, batch: int = 64, seed: int = 0):
        rng = np.random.default_rng(seed)
        N = X.shape[0]
        losses = []
        for ep in range(epochs):
            idx = rng.permutation(N)[: min(batch, N)]
            loss, g = self.gradients(X[idx], T[idx], M[idx])
            self.step += 1
            self.cv.lib.cvlhip_mlp_adam(self.params, g, self.m, self.v, self.step, lr)
            losses.append(loss)
        return losses

    def choose(self, feat: np.ndarray) -> int:
        return int(masked_argmin(self.predict(feat[None, :]), admitted(feat))[0])

    def install(self):
        """Make the library's mke_fov pick with these parameters."""
        self.cv.lib.cvlhip_sched_set_params(self.params)

    def save(self, path: str = MODEL_PATH):
        np.savez(path, params=self.params.cpu().numpy(),
                 dims=np.array([self.F, self.H, self.A, self.P], np.int64))

    def load(self, path: str = MODEL_PATH) -> "Dispatcher":
        with np.load(path, allow_pickle=False) as z:
            if "dims" in z and tuple(int(v) for v in z["dims"]) != (self.F, self.H, self.A, self.P):
                raise ValueError(f"{path}: network shape {tuple(z['dims'])} does not match the library's "
                                 f"{(self.F, self.H, self.A, self.P)}")
            self.params = self._dev(z["params"])
        return self


# ------------------------------------------------------------------ data
class Inputs:
    """Value vectors of the largest size, both types, from splitmix64."""

    def __init__(self, cv, nmax: int):
        import torch
        self.xi = torch.empty(nmax, dtype=torch.int32, device=cv.device)
        cv.lib.cvlhip_fill_splitmix_u32(self.xi, nmax, 0x4E45534C00000003)
        self.xd = torch.empty(nmax, dtype=torch.float64, device=cv.device)
        cv.lib.cvlhip_fill_splitmix_f64(self.xd, nmax, 0x4E45534C00000003, -1.0, 1.0)
        self.out_i = torch.empty_like(self.xi)
        self.out_d = torch.empty_like(self.xd)

    def of(self, ty, n):
        return (self.xd[:n], self.out_d[:n]) if ty == "d" else (self.xi[:n], self.out_i[:n])


def gather(cv, dists, sizes, kinds=KINDS, seed: int = 0, log=print, reps: int = 5):
    """(X, T, M, meta): features and log-times of every arm for every
    (distribution, size, kind)."""
    tuner = Autotuner(cv, reps)
    rng = np.random.default_rng(seed)
    inp = Inputs(cv, max(sizes))
    X, T, M, meta = [], [], [], []
    for (dist, param) in dists:
        for n in sizes:
            lens = lengths(dist, n, param, rng)
            sd = cv.segd(lens)
            for (k, op, ty) in kinds:
                s, d = inp.of(ty, n)
                t = tuner.tune(k, op, ty, d, s, sd)
                X.append(features(cv, sd, k, ty))
                T.append(np.log(t))
                M.append(np.ones(len(ARMS), np.float32))
                meta.append({"dist": dist, "param": param, "n": n, "m": int(lens.size), "op": f"{op}_{k}_{ty}",
                             "ms": [round(float(v), 5) for v in t]})
                log(json.dumps(meta[-1]))
    return (np.array(X, np.float32), np.array(T, np.float32), np.array(M, np.float32), meta)


def offline_regret(disp: Dispatcher, X, T):
    """Regret of the network's picks on measured arm times (no decision cost)."""
    times = np.exp(T)
    pick = masked_argmin(disp.predict(X), admitted(X))
    chosen = times[np.arange(len(X)), pick]
    return float(chosen.sum() / times.min(axis=1).sum() - 1.0)


# ------------------------------------------------------------------ live evaluation
def evaluate_live(cv, dists, sizes, kinds=KINDS, seed: int = 11, reps: int = 7, log=print):
    """Times, per (distribution, size): every arm of every kind (the
    autotuner's table), the CVL entry (add_sez, add_sed, max_rez, max_red)
    on a descriptor built by mke_fov with the library's dispatcher on, and
    the decision itself: mke_fov with the dispatcher on minus off (the
    histogram and the network's forward pass), charged once per descriptor.
    Returns the per-case records and the totals."""
    import torch
    from .cvl import Segd
    L = cv.lib
    timer = Timer(reps)
    tuner = Autotuner(cv, reps)
    rng = np.random.default_rng(seed)
    inp = Inputs(cv, max(sizes))
    cases, tot_arm, tot_learned, tot_best, tot_decision = [], np.zeros(len(ARMS)), 0.0, 0.0, 0.0
    for (dist, param) in dists:
        for n in sizes:
            lens_h = lengths(dist, n, param, rng)
            m = int(lens_h.size)
            lens = torch.from_numpy(lens_h).to(cv.device)
            units = L.siz_fos(n, m)
            bufs = {on: torch.zeros(units * 8 // 4 + 64, dtype=torch.int32, device=cv.device) for on in (0, 1)}
            bscr = torch.empty(L.mke_fov_scratch(n, m) * 8 + 256, dtype=torch.uint8, device=cv.device)
            build_ms = {}
            for on in (0, 1):
                L.cvlhip_sched_enable(on)
                build_ms[on] = timer(lambda on=on: L.mke_fov(bufs[on], lens, n, m, bscr))
            L.cvlhip_sched_enable(1)
            decision = max(0.0, build_ms[1] - build_ms[0])
            sd = Segd(bufs[1], n, m)
            rec = {"dist": dist, "param": param, "n": n, "m": m, "build_ms": build_ms[0],
                   "decision_ms": round(decision, 5), "kinds": {}}
            for kind in kinds:
                k, op, ty = kind
                s, d = inp.of(ty, n)
                arms = tuner.tune(k, op, ty, d if k == "scan" else d[:m], s, sd)
                name = CVL_NAME[kind]
                scr = cv.scratch(getattr(L, name + "_scratch")(n, m))
                dst = d if k == "scan" else d[:m]
                entry = timer(lambda: getattr(L, name)(dst, s, sd, n, m, scr))
                pick = L.cvlhip_segd_pick(sd, 1 if k == "reduce" else 0, 8 if ty == "d" else 4)
                rec["kinds"][name] = {"arm_ms": [round(float(v), 5) for v in arms],
                                      "best_arm": ARMS[int(arms.argmin())], "pick": ARMS[pick] if pick >= 0 else None,
                                      "entry_ms": round(entry, 5)}
                tot_arm += arms
                tot_best += float(arms.min())
                tot_learned += entry
            tot_learned += decision
            tot_decision += decision
            cases.append(rec)
            log(json.dumps(rec))
    fixed = {ARMS[a]: float(tot_arm[a]) for a in range(len(ARMS))}
    best_fixed = min(fixed, key=fixed.get)
    return {
        "cases": cases,
        "learned_total_ms": tot_learned,
        "decision_total_ms": tot_decision,
        "autotuned_total_ms": tot_best,
        "regret": tot_learned / tot_best - 1.0,
        "fixed_arm_total_ms": fixed,
        "best_fixed_arm": best_fixed,
        "learned_vs_best_fixed": tot_learned / fixed[best_fixed] - 1.0,
    }


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    from .cvl import Cvl
    cv = Cvl()
    t0 = time.time()
    if argv and argv[0] == "eval":
        # the library's built-in network on the held-out workloads
        if not cv.lib.cvlhip_sched_has_model():
            raise SystemExit("libcvl_hip.so was built without a dispatcher model")
        cv.lib.cvlhip_sched_enable(1)
        res = {"model": "built into libcvl_hip.so from dispatch_model.npz",
               "eval": evaluate_live(cv, EVAL_DISTS, EVAL_SIZES), "seconds": round(time.time() - t0, 1)}
        if len(argv) > 1:
            with open(argv[1], "w") as f:
                json.dump(res, f, indent=1)
        print(json.dumps({k: v for k, v in res["eval"].items() if k != "cases"}), flush=True)
        return
    if argv and argv[0] != "train":
        raise SystemExit(__doc__)
    X, T, M, meta = gather(cv, TRAIN_DISTS, TRAIN_SIZES, seed=0)
    Xv, Tv, Mv, metav = gather(cv, VALID_DISTS, VALID_SIZES, seed=5)
    # the training run: the seed and length with the lowest validation regret
    best = None
    for seed in (0, 1, 2):
        disp = Dispatcher(cv, seed=seed)
        done = 0
        for epochs in (1500, 1500, 3000):
            losses = disp.train(X, T, M, epochs=epochs, seed=seed + done)
            done += epochs
            r = offline_regret(disp, Xv, Tv)
            print(json.dumps({"seed": seed, "epochs": done, "loss": float(np.mean(losses[-50:])),
                              "valid_regret": r}), flush=True)
            if best is None or r < best[0]:
                best = (r, seed, done, disp.params.clone())
    disp = Dispatcher(cv, seed=best[1])
    disp.params = best[3]
    disp.save(argv[2] if len(argv) > 2 else MODEL_PATH)
    disp.install()
    res = {"train_cases": int(len(X)), "valid_cases": int(len(Xv)), "valid_regret": best[0],
           "chosen_seed": best[1], "chosen_epochs": best[2],
           "train_offline_regret": offline_regret(disp, X, T),
           "eval": evaluate_live(cv, EVAL_DISTS, EVAL_SIZES), "seconds": round(time.time() - t0, 1),
           "train_timings": meta, "valid_timings": metav}
    print(json.dumps({k: v for k, v in res.items() if k not in ("train_timings", "valid_timings")})[:4000],
          flush=True)
    out = argv[1] if len(argv) > 1 else None
    if out:
        with open(out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
