"""Benchmark: the CVL primitive sweep at 2^28 int32 on one MI355X.

One step = one pass of the sweep, each primitive called once through the
C-ABI of libcvl_hip.so on inputs already resident in HBM:
    add_suz   unsegmented exclusive +-scan, 2^28 int32           8 B/elem
    max_ruz   unsegmented max-reduce, 2^28 int32                 4 B/elem
    add_sez   segmented +-scan, geometric lengths (mean 256)     8 + 4m/n B/elem
    pk2_luz   pack, Bernoulli(0.5) flags                         4 + 4 + 4*kept/n B/elem
    smp_puz   permute by a random permutation                    12 B/elem
    rku_luz   radix rank of 2^26 uniform 32-bit keys             8 B/key
The byte counts are compulsory traffic (each input read once, each output
written once, at CVL widths).  `value` is the whole sweep's compulsory bytes
divided by its time (GB/s); per-primitive figures are in "primitives".

    python bench.py [--gpus 1] [--steps 5] [--warmup 2]

Prints one JSON line.  The reference serial CVL (compiled from its sources
into oracle/_ref, else the C oracle) is timed on the host on a bounded
sample of the same workload as `cpu_baseline`.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 0x4E45534C00000000
CONFIG_ID = 2


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--log2n", type=int, default=28)
    p.add_argument("--rank-log2n", type=int, default=26)
    p.add_argument("--no-cpu", action="store_true", help="skip the host baseline")
    p.add_argument("--only", default="", help="comma list of primitives to run (profiling)")
    p.add_argument("--no-irregular", action="store_true",
                   help="skip the irregular float64 segmented workload and its dispatch comparison")
    p.add_argument("--irregular-log2n", type=int, default=26)
    p.add_argument("--no-programs", action="store_true",
                   help="skip the whole-program workloads (quicksort, quickhull, SpMV through vinterp)")
    p.add_argument("--programs", default="",
                   help="comma list of program workloads to run alone (profiling), e.g. quickhull_fused,spmv_fused")
    return p.parse_args()


from neslgpu_amd.workloads import geometric_lengths  # noqa: E402


class Sweep:
    """Device inputs for the primitives in `names` (only those are built)."""

    def __init__(self, log2n: int, rank_log2n: int, names):
        import torch
        from neslgpu_amd.cvl import Cvl
        self.torch = torch
        self.cv = Cvl()
        self.lib = self.cv.lib
        n = 1 << log2n
        nr = 1 << rank_log2n
        self.n, self.nr = n, nr
        self.m, self.kept = 0, 0
        dev = self.cv.device
        i32 = torch.int32
        L = self.lib
        # inputs: splitmix64 streams generated on the device
        self.s = torch.empty(n, dtype=i32, device=dev)
        L.cvlhip_fill_splitmix_u32(self.s, n, SEED + CONFIG_ID)
        self.d = torch.empty(n, dtype=i32, device=dev)
        units = [1, L.add_suz_scratch(n), L.max_ruz_scratch(n)]
        if "rku_luz" in names:
            self.keys = torch.empty(nr, dtype=i32, device=dev)
            L.cvlhip_fill_splitmix_u32(self.keys, nr, SEED + CONFIG_ID + 1)
            self.rank = torch.empty(nr, dtype=i32, device=dev)
            units.append(L.rku_luz_scratch(nr))
        if "pk2_luz" in names:
            u = torch.empty(n, dtype=i32, device=dev)
            L.cvlhip_fill_splitmix_u32(u, n, SEED + CONFIG_ID + 2)
            self.flags = (u < 0).to(i32)                   # top bit: Bernoulli(0.5)
            del u
            self.kept = int(self.flags.sum().item())
            units.append(L.pk2_luz_scratch(n))
        if "smp_puz" in names:
            g = torch.Generator(device=dev)
            g.manual_seed(SEED + CONFIG_ID + 3 & 0x7FFFFFFF)
            self.perm = torch.randperm(n, device=dev, generator=g, dtype=torch.int64).to(i32)
            units.append(L.smp_puz_scratch(n))
        if "add_sez" in names:
            self.lens = geometric_lengths(n, 256, SEED + CONFIG_ID + 4 & 0x7FFFFFFF)
            self.m = int(self.lens.size)
            self.segd = self.cv.segd(self.lens)
            units.append(L.add_sez_scratch(n, self.m))
        # one scratch area serves every primitive: the largest of their
        # *_scratch() requirements (CVL units of 8 bytes)
        self.scr = torch.empty(max(units) * 8 + 4096, dtype=torch.uint8, device=dev)
        self.max_result = None
        torch.cuda.synchronize()

    # compulsory bytes of each primitive
    def bytes(self) -> dict:
        n, nr, m = self.n, self.nr, self.m
        return {
            "add_suz": 8 * n,
            "max_ruz": 4 * n,
            "add_sez": 8 * n + 4 * m,
            "pk2_luz": 8 * n + 4 * self.kept,
            "smp_puz": 12 * n,
            "rku_luz": 8 * nr,
            "pack_full": 12 * n + 4 * self.kept,
            "cpy_wuz": 8 * n,
        }

    def run(self, name: str):
        L, n = self.lib, self.n
        if name == "add_suz":
            L.add_suz(self.d, self.s, n, self.scr)
        elif name == "max_ruz":
            self.max_result = L.max_ruz(self.s, n, self.scr)
        elif name == "add_sez":
            L.add_sez(self.d, self.s, self.segd, n, self.m, self.scr)
        elif name == "pk2_luz":
            L.pk2_luz(self.d, self.s, self.flags, n, self.kept, self.scr)
        elif name == "smp_puz":
            L.smp_puz(self.d, self.s, self.perm, n, self.scr)
        elif name == "rku_luz":
            L.rku_luz(self.rank, self.keys, self.nr, self.scr)
        elif name == "pack_full":          # VCODE PACK: count (host read), then compact
            kept = L.pk1_luv(self.flags, n, self.scr)
            L.pk2_luz(self.d, self.s, self.flags, n, kept, self.scr)
        elif name == "cpy_wuz":            # calibration copy (hipMemcpy device to device)
            L.cpy_wuz(self.d, self.s, n, self.scr)

    def step(self, names):
        for k in names:
            self.run(k)


def time_gpu(fn, reps: int, torch) -> float:
    """Average milliseconds of fn() over reps, HIP events on the current stream."""
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


# ------------------------------------------------------------------ irregular float64 work
def irregular(log2n: int, reps: int) -> dict:
    """Segmented float64 +-scan and max-reduce over Zipf(1.5) segment lengths
    clipped to [1, 2^20] (2^26 values uniform in [-1, 1]): every schedule
    timed (median of `reps`, HIP events), the exhaustive autotuner's best,
    and the learned dispatch as a program gets it: the CVL entry points
    add_sed / max_red on a descriptor whose schedules mke_fov picked with the
    library's network, plus the cost of that decision (mke_fov with the
    dispatcher on minus off).  Also both operations in one pass
    (cvlhip_add_sed_max_red)."""
    import torch
    from neslgpu_amd import dispatch as D
    from neslgpu_amd.cvl import Cvl, Segd
    cv = Cvl()
    L = cv.lib
    n = 1 << log2n
    x = torch.empty(n, dtype=torch.float64, device=cv.device)
    L.cvlhip_fill_splitmix_f64(x, n, SEED + 3, -1.0, 1.0)
    lens = D.lengths("zipf", n, 1.5, np.random.default_rng(SEED + 3 & 0x7FFFFFFF), lmax=1 << 20)
    m = int(lens.size)
    lens_dev = torch.from_numpy(lens).to(cv.device)
    units = L.siz_fos(n, m)
    bufs = {on: torch.zeros(units * 2 + 64, dtype=torch.int32, device=cv.device) for on in (0, 1)}
    bscr = torch.empty(L.mke_fov_scratch(n, m) * 8 + 256, dtype=torch.uint8, device=cv.device)
    build = {}
    for on in (0, 1):
        L.cvlhip_sched_enable(on)
        build[on] = float(np.median([time_gpu(lambda on=on: L.mke_fov(bufs[on], lens_dev, n, m, bscr), 1, torch)
                                     for _ in range(reps)]))
    L.cvlhip_sched_enable(1)
    decision = max(0.0, build[1] - build[0])
    sd = Segd(bufs[1], n, m)
    d = torch.empty_like(x)
    r = torch.empty(m, dtype=torch.float64, device=cv.device)
    tuner = D.Autotuner(cv, reps=reps)
    scr = cv.scratch(max(L.add_sed_scratch(n, m), L.max_red_scratch(n, m), L.cvlhip_seg_scratch(n, m),
                         L.cvlhip_add_sed_max_red_scratch(n, m)))
    out = {"workload": f"2^{log2n} float64 uniform [-1,1], Zipf(1.5) segment lengths clipped to [1,2^20]",
           "n": n, "segments": m, "arms": D.ARMS, "model_built_in": bool(L.cvlhip_sched_has_model()),
           "mke_fov_ms": {"dispatcher_off": round(build[0], 4), "dispatcher_on": round(build[1], 4)},
           "decision_ms": round(decision, 4)}
    tot = {"learned": decision, "autotuned": 0.0}
    fixed = np.zeros(len(D.ARMS))
    for kind, op, dst, byts in (("scan", "add", d, 16 * n + 4 * m), ("reduce", "max", r, 8 * n + 12 * m)):
        t = tuner.tune(kind, op, "d", dst, x, sd)
        fixed += t
        best = int(np.argmin(t))
        pick = L.cvlhip_segd_pick(sd, 1 if kind == "reduce" else 0, 8)
        if kind == "scan":
            entry = lambda: L.add_sed(d, x, sd, n, m, scr)
        else:
            entry = lambda: L.max_red(r, x, sd, n, m, scr)
        te = float(np.median([time_gpu(entry, 1, torch) for _ in range(reps)]))
        rec = {"ms_per_arm": [round(float(v), 4) for v in t], "autotuned_arm": D.ARMS[best],
               "autotuned_ms": round(float(t[best]), 4), "learned_arm": D.ARMS[pick] if pick >= 0 else None,
               "learned_ms": round(te, 4), "compulsory_bytes": byts,
               "learned_GB/s": round(byts / te / 1e6, 1)}
        tot["autotuned"] += float(t[best])
        tot["learned"] += te
        out[f"{op}_{kind}"] = rec
    # both operations in one pass over the values (cvlhip_add_sed_max_red)
    pair = lambda: L.cvlhip_add_sed_max_red(d, r, x, sd, n, m, scr)
    tp = float(np.median([time_gpu(pair, 1, torch) for _ in range(reps)]))
    tot["one_pass_pair"] = tp
    out["one_pass_pair"] = {"ms": round(tp, 4), "compulsory_bytes": 16 * n + 4 * m + 8 * m,
                            "GB/s": round((16 * n + 12 * m) / tp / 1e6, 1)}
    out["total_ms"] = {k: round(v, 4) for k, v in tot.items()}
    out["fixed_arm_total_ms"] = {D.ARMS[a]: round(float(fixed[a]), 4) for a in range(len(D.ARMS))}
    out["learned_regret_vs_autotuned"] = round(tot["learned"] / tot["autotuned"] - 1.0, 4)
    out["learned_vs_best_fixed"] = round(tot["learned"] / float(fixed.min()) - 1.0, 4)
    return out


# ------------------------------------------------------------------ whole programs
PROGRAMS = os.path.join(ROOT, "tests", "golden", "vcode", "programs")
OWN_PROGRAMS = os.path.join(ROOT, "neslgpu_amd", "vcode", "programs")


def program_workloads():
    """(name, program text, stdin, fused library or None, graphs, repeats,
    description) of the whole-program workloads."""
    from neslgpu_amd import vinterp as V
    from neslgpu_amd.build import generated
    from neslgpu_amd import workloads as W
    qs = W.qs_bench_program(V.read(os.path.join(PROGRAMS, "qs.1000.vcode")))
    qsf = V.read(generated("qs_bench", "fcode"))          # the same program through this package's fuser
    qh = V.set_size(V.read(os.path.join(PROGRAMS, "quickhull.1000.vcode")), 1 << 24)
    qhf = V.set_size(V.read(generated("quickhull.1000", "fcode")), 1 << 24)
    rows = W.spmv_rows()
    spmv_in = W.spmv_stdin(rows, W.SPMV_ROWS)
    spmv = V.read(os.path.join(OWN_PROGRAMS, "spmv.vcode"))
    spmvf = V.read(generated("spmv", "fcode"))
    return [
        ("quicksort", qs, None, None, False, 1, "qs.nesl: 2^20 keys uniform in [0, 2^31-1) (RAND), sorted"),
        ("quicksort_fused", qsf, None, generated("qs_bench", "so"), False, 1,
         "same, fused by neslgpu_amd.fusion.fuser (FUSED super-ops, generated HIP)"),
        ("quickhull", qh, None, None, False, 1, "quickhull.nesl: 2^24 points uniform in the unit disk"),
        ("quickhull_fused", qhf, None, generated("quickhull.1000", "so"), False, 1,
         "same, fused by neslgpu_amd.fusion.fuser (FUSED super-ops, generated HIP)"),
        ("spmv", spmv, spmv_in, None, False, 10, "y = A x, CSR 2^22 rows, power-law lengths (mean 16, "
         "max 2^16), 2^26 nnz, float64; BPERMUTE, *, +_REDUCE"),
        ("spmv_fused", spmvf, spmv_in, generated("spmv", "so"), False, 10,
         "same, fused by neslgpu_amd.fusion.fuser: gather, multiply and segmented sum in one kernel"),
        ("spmv_fused_graph", spmvf, spmv_in, generated("spmv", "so"), True, 10,
         "same, fused block replayed as a HIP graph"),
    ], rows


def program_timer(stdout: str) -> float:
    return float(stdout.splitlines()[-1].strip("() "))


def programs(reps: int, only=()) -> dict:
    """Whole VCODE programs through vinterp on libcvl_hip.so: median of the
    programs' own START_TIMER / STOP_TIMER interval (data generation
    excluded), in ms.  SpMV repeats the product 10 times inside the timed
    interval; its figures are per product.  `only`: names to run (all when
    empty)."""
    from neslgpu_amd import vinterp as V
    work, rows = program_workloads()
    if only:
        work = [w for w in work if w[0] in only]
    out = {}
    nnz, m, ncols = int(rows.sum()), int(rows.size), 1 << 22
    for name, text, stdin, fused, graphs, k, desc in work:
        ts, first = [], None
        for _ in range(reps):
            r = V.run(text, fused=fused, graphs=graphs, stdin=stdin, arena_units=(1 << 31) - 1, timeout=300)
            ts.append(program_timer(r.stdout) / k * 1e3)
            first = first or r.stdout.splitlines()[:-1]
        rec = {"ms": round(float(np.median(ts)), 4), "runs_ms": [round(t, 4) for t in ts], "workload": desc,
               "output": first}
        if name.startswith("spmv"):
            byts = nnz * 12 + ncols * 8 + m * (4 + 8)
            rec["ms_per_product"] = rec.pop("ms")
            rec["compulsory_bytes"] = byts
            rec["GB/s"] = round(byts / (rec["ms_per_product"] * 1e-3) / 1e9, 1)
            rec["sum_y"] = first[-1] if first else None
        out[name] = rec
    return out


def cpu_programs() -> dict:
    """The same programs through vinterp on the serial reference CVL
    (oracle/_ref), single thread: one run each, ms."""
    from neslgpu_amd import vinterp as V
    from oracle import refcvl
    if not refcvl.available():
        return {"skipped": "reference serial CVL not built"}
    serial = refcvl.LIB_PATH
    work, _ = program_workloads()
    out = {}
    for name, text, stdin, fused, graphs, k, desc in work:
        if fused:
            continue                              # fused libraries are HIP code
        r = V.run(text, backend=serial, stdin=stdin, arena_units=(1 << 31) - 1, timeout=600)
        out[name] = {"ms" if k == 1 else "ms_per_product": round(program_timer(r.stdout) / k * 1e3, 3),
                     "output": r.stdout.splitlines()[:-1]}
    return out


def compare_outputs(gpu: dict, cpu: dict) -> None:
    """Marks each GPU program record with matches_serial: the values the
    program printed (its timer aside) against the serial CVL's in the same
    run -- equal text for integer results (quicksort's keys, quickhull's
    hull), and for SpMV's sum of y a relative difference of at most 1e-12
    (the sums differ in association only).  The output text itself is
    dropped from the record."""
    for name, rec in gpu.items():
        got = rec.pop("output", None)
        ref = cpu.get(name.replace("_fused_graph", "").replace("_fused", ""), {}).get("output")
        if got is None or ref is None:
            continue
        if name.startswith("spmv"):
            a, b = float(got[-1].strip("() ")), float(ref[-1].strip("() "))
            rec["serial_sum_y"] = ref[-1]
            rec["matches_serial"] = abs(a - b) <= 1e-12 * max(1.0, abs(b))
        else:
            rec["matches_serial"] = got == ref
    for rec in cpu.values():
        if isinstance(rec, dict):
            rec.pop("output", None)


def measured_traffic(primitive: str):
    """HBM bytes per call of `primitive` from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel
    calibration), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    rec = t.get("primitives", {}).get(primitive)
    if not rec:
        return None
    return rec.get("hbm_bytes")


# ------------------------------------------------------------------ host baseline
def cpu_baseline(budget_s: float = 20.0) -> dict:
    """The serial CVL on this host, single thread, on a bounded sample of the
    sweep: 2^24 elements for the n-sized primitives (64 MiB per vector, far
    beyond the host caches), 2^22 keys for rank."""
    from oracle import oracle as O
    from oracle import refcvl
    kind = "reference" if refcvl.available() else "port"
    n = 1 << 24
    nr = 1 << 22
    s = O.splitmix_u32(n, SEED + CONFIG_ID).view(np.int32)
    keys = O.splitmix_u32(nr, SEED + CONFIG_ID + 1).view(np.int32)
    flags = (O.splitmix_u32(n, SEED + CONFIG_ID + 2) >> 31).astype(np.int32)
    rng = np.random.default_rng(1)
    perm = rng.permutation(n).astype(np.int32)
    lens = geometric_lengths(n, 256, SEED + CONFIG_ID + 4 & 0x7FFFFFFF)
    m = lens.size
    kept = int(flags.sum())
    d = np.empty(n, np.int32)
    rk = np.empty(nr, np.int32)
    scr = np.zeros(4 * nr + 16, np.float64)
    byts = {"add_suz": 8 * n, "max_ruz": 4 * n, "add_sez": 8 * n + 4 * m, "pk2_luz": 8 * n + 4 * kept,
            "smp_puz": 12 * n, "rku_luz": 8 * nr}
    times = {}
    t_start = time.perf_counter()
    if kind == "reference":
        call = refcvl.call
        ops = {
            "add_suz": lambda: call("add_suz", d, s, n, scr),
            "max_ruz": lambda: call("max_ruz", s, n, scr),
            "add_sez": lambda: call("add_sez", d, s, lens, n, m, scr),
            "pk2_luz": lambda: call("pk2_luz", d, s, flags, n, kept, scr),
            "smp_puz": lambda: call("smp_puz", d, s, perm, n, scr),
            "rku_luz": lambda: call("rku_luz", rk, keys, nr, scr),
        }
    else:
        ops = {
            "add_suz": lambda: O.scan("add", "z", s),
            "max_ruz": lambda: O.reduce("max", "z", s),
            "add_sez": lambda: O.segscan("add", "z", s, lens),
            "pk2_luz": lambda: O.pk2("z", s, flags),
            "smp_puz": lambda: O.smp_pu("z", s, perm),
            "rku_luz": lambda: O.rank(keys, True),
        }
    for k, f in ops.items():
        f()                                  # warm caches and page in
        reps = 0
        t0 = time.perf_counter()
        while True:
            f()
            reps += 1
            el = time.perf_counter() - t0
            if el > budget_s / 12 or reps >= 5:
                break
        times[k] = el / reps
    total_bytes = sum(byts.values())
    total_time = sum(times.values())
    return {
        "value": round(total_bytes / total_time / 1e9, 3),
        "unit": "GB/s",
        "cores": 1,
        "kind": kind,
        "sample": "same sweep at 2^24 elements (rank: 2^22 keys), splitmix64 inputs, serial CVL -O2 "
                  "-fwrapv, mean of up to 5 timed runs per primitive after one warm-up run",
        "per_primitive_gbs": {k: round(byts[k] / times[k] / 1e9, 3) for k in times},
        "host_cpu": _cpu_model(),
        "host_nproc": os.cpu_count(),
        "wall_s": round(time.perf_counter() - t_start, 2),
    }


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def progress(msg: str) -> None:
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported "
              "(this build has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    names = ["add_suz", "max_ruz", "add_sez", "pk2_luz", "smp_puz", "rku_luz"]
    if a.only:
        names = [x for x in names + ["cpy_wuz"] if x in a.only.split(",")]
    sw = Sweep(a.log2n, a.rank_log2n, names)
    byts = sw.bytes()
    progress("inputs ready")
    for _ in range(a.warmup):
        sw.step(names)
    torch.cuda.synchronize()
    progress("warm-up done")
    t0 = time.perf_counter()
    ms = time_gpu(lambda: sw.step(names), a.steps, torch)
    wall = (time.perf_counter() - t0) / a.steps * 1e3
    total_bytes = sum(byts[k] for k in names)
    prims = {}
    for k in names:
        t = time_gpu(lambda k=k: sw.run(k), a.steps, torch)
        gbs = byts[k] / t / 1e6
        prims[k] = {"ms": round(t, 4), "GB/s": round(gbs, 1), "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
                    "bytes": byts[k]}
        progress(f"{k} {prims[k]['ms']} ms")
    dom = max(prims, key=lambda k: prims[k]["ms"])
    extra = {}
    if "pk2_luz" in names:
        # the whole VCODE PACK (ref nesl/vcode/main.c:577-642): pk1_luv
        # counts the flags and returns the count to the host, then pk2_luz
        t = time_gpu(lambda: sw.run("pack_full"), a.steps, torch)
        gbs = byts["pack_full"] / t / 1e6
        extra["pack_full"] = {"ms": round(t, 4), "GB/s": round(gbs, 1), "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
                              "bytes": byts["pack_full"], "calls": "pk1_luv + pk2_luz (count read by the host)"}
    kernel_of = {"add_suz": "k_scan (AddZ, unsegmented)", "max_ruz": "k_reduce_strided + k_reduce_finish (MaxZ)",
                 "add_sez": "k_scan (AddZ, segmented)", "pk2_luz": "k_pack<int>",
                 "smp_puz": "k_part_hist + k_tile_scan + k_partition<int> x2 (pairs, then 16-bit offsets) "
                            "+ k_block_scatter<int>",
                 "rku_luz": "k_key_hist + 4 x (k_rank_hist + k_tile_scan + k_rank_pass) + "
                            "inversion (k_part_hist + k_tile_scan + k_partition<int> x2 + k_block_scatter<int>)",
                 "cpy_wuz": "__amd_rocclr_copyBuffer"}
    roof = {
        "bound": "hbm",
        "kernel": kernel_of[dom],
        "primitive": dom,
        "achieved": prims[dom]["GB/s"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": prims[dom]["frac_of_hbm_peak"],
        "traffic": measured_traffic(dom),
        "traffic_source": "profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)",
    }
    value = total_bytes / ms / 1e6
    out = {
        "metric": "CVL primitive GB/s vs HBM peak; quickhull/SpMV time, learned vs fixed dispatch",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 4),
        "wall_ms_per_step": round(wall, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (splitmix64 keys/values, geometric(mean 256) segment lengths, "
                "Bernoulli(0.5) flags, uniform random permutation)",
        "config": {"workload": f"CVL primitive sweep: 2^{a.log2n} int32 (add_suz, max_ruz, add_sez, "
                               f"pk2_luz, smp_puz) + rku_luz on 2^{a.rank_log2n} keys",
                   "n": sw.n, "segments": sw.m, "rank_keys": sw.nr, "primitives_run": names},
        "frac_of_hbm_peak": round(value / HBM_PEAK_GBS, 4),
        "primitives": prims,
        "outside_sweep": extra,
        "roofline": roof,
    }
    if not a.no_irregular and not a.only:
        out["irregular_f64"] = irregular(a.irregular_log2n, max(3, a.steps))
        progress("irregular done")
    if a.programs:
        out["programs"] = programs(3, tuple(a.programs.split(",")))
    elif not a.no_programs and not a.only:
        out["programs"] = programs(3)
    progress("programs done")
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        if not a.no_programs and not a.only and not a.programs:
            out["cpu_baseline"]["programs"] = cpu_programs()
            compare_outputs(out["programs"], out["cpu_baseline"]["programs"])
    for rec in out.get("programs", {}).values():
        rec.pop("output", None)
    out["children_at_exit"] = children()
    print(json.dumps(out), flush=True)


def children() -> list:
    """Command lines of this process's live descendants (bench.py starts
    only vinterp runs and waits for each; any listed here is reported, not
    left unexplained)."""
    try:
        import psutil
        procs = psutil.Process().children(recursive=True)
        return [" ".join(p.cmdline())[:200] for p in procs if p.is_running()]
    except Exception as e:          # psutil missing or a process gone meanwhile
        return [f"unknown: {e!r}"]


if __name__ == "__main__":
    main()
