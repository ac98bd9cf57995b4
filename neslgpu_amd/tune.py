"""Autotuner for the kernels' compile-time tile shapes.

The HIP sources take their tile and workgroup shapes from CVLHIP_* macros
(`#ifndef CVLHIP_X / #define CVLHIP_X default`).  This tool searches them
one at a time around the current values (coordinate search: each candidate
of one macro, every other macro at its current value):

    python -m neslgpu_amd.tune build [KEY ...]
        # on the build host: every variant library (of the named SPACE keys,
        # or of all), compiled with -D for the
        # one macro (only the sources that see the macro are recompiled,
        # the other objects are the default build's), into neslgpu_amd/_tune/
    python -m neslgpu_amd.tune run [out.json]
        # on the GPU: each variant in a process of its own (a variant is a
        # whole libcvl_hip.so): its primitive checked against the oracle on
        # a small input, then timed at the benchmark's size
    python -m neslgpu_amd.tune pick out.json
        # the fastest correct value of each macro -> neslgpu_amd/csrc/tuned.hpp
        # (build.py force-includes it into every HIP source and fused
        # library) and the record -> profiles/tile_autotune.md

A variant that does not compile (a shape past the LDS or register budget,
or a static_assert) is left out; one whose results differ from the oracle's
is never picked.
"""
from __future__ import annotations

import glob
import json
import os
import re
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
TUNE_DIR = os.path.join(HERE, "_tune")
TUNED = os.path.join(CSRC, "tuned.hpp")
RECORD = os.path.join(ROOT, "profiles", "tile_autotune.md")

# macro -> (candidate values, the primitive whose time judges it)
SPACE = {
    "CVLHIP_SCAN_ROWS_U4,CVLHIP_SCAN_BLOCK_U4": ([(12, 1024), (8, 1024), (8, 512), (12, 512), (16, 512), (16, 256)],
                                                 "add_suz"),
    "CVLHIP_LOOKBACK_K": ([1, 2, 4], "add_suz"),
    "CVLHIP_SEG_SCAN_BLOCK4": ([256, 512, 1024], "add_sez"),
    "CVLHIP_SEG_SCAN_BLOCK8": ([1024, 512, 256], "add_sed"),
    "CVLHIP_REDUCE_GRID": ([1024, 2048], "max_ruz"),
    "CVLHIP_PACK_ROWS": ([4, 6, 8, 10], "pk2_luz"),
    "CVLHIP_PART_ITEMS": ([4, 8, 12], "smp_puz"),
    "CVLHIP_PART_THREADS": ([512, 1024], "smp_puz"),
    "CVLHIP_BLK_LOG4": ([12, 13, 14], "smp_puz"),
    "CVLHIP_PART_ATOMIC_FIRST": ([0, 1], "smp_puz"),
    "CVLHIP_RANK_THREADS,CVLHIP_RANK_ITEMS": ([(512, 8), (256, 16), (512, 16), (1024, 8), (256, 8), (1024, 4)],
                                              "rku_luz"),
    "CVLHIP_RANK_CHAIN": ([0, 1], "rku_luz"),
    "CVLHIP_TILE_SCAN_ROWS": ([64, 32, 128, 16], "rku_luz"),
    "CVLHIP_SEG_REDUCE_ROWS,CVLHIP_SEG_REDUCE_BLOCK,CVLHIP_SEG_REDUCE_LANE_BYTES": (
        [(2, 256, 64), (8, 256, 16), (4, 256, 32), (4, 256, 64), (1, 256, 64), (2, 512, 64), (3, 256, 64)],
        "max_red"),
    # a comma-joined key searches its macros together (values are tuples)
    "CVLHIP_PAIR_ROWS,CVLHIP_PAIR_BLOCK,CVLHIP_PAIR_LANE_BYTES": (
        [(8, 512, 16), (4, 512, 32), (2, 512, 64), (2, 256, 64), (4, 256, 64), (1, 512, 64), (2, 1024, 64)],
        "add_sed_max_red"),
}


def macros(key: str) -> list:
    return key.split(",")


def key_value(key: str):
    """Current value of a SPACE key (a tuple for a joined key)."""
    vals = tuple(current_value(m)[0] for m in macros(key))
    return vals if len(vals) > 1 else vals[0]


def norm(v):
    """Values compare as tuples (JSON gives joined values back as lists)."""
    return tuple(v) if isinstance(v, (list, tuple)) else v


def defines(key: str, value) -> list:
    vals = value if isinstance(value, (tuple, list)) else (value,)
    return [f"-D{m}={v}" for m, v in zip(macros(key), vals)]


# ------------------------------------------------------------------ build host
def current_value(macro: str):
    """(default value, file that defines it) of a macro, from the sources
    (or tuned.hpp when it sets the macro)."""
    if os.path.exists(TUNED):
        m = re.search(rf"#define {macro} (\S+)", open(TUNED).read())
        if m:
            return int(m.group(1)), None
    for f in glob.glob(os.path.join(CSRC, "*.h*")):
        if f == TUNED:
            continue
        m = re.search(rf"#ifndef {macro}\s*\n#define {macro} (\S+)", open(f).read())
        if m:
            return int(m.group(1)), f
    raise KeyError(macro)


def includers(path: str) -> list:
    """The .hip sources that include `path`, directly or through headers."""
    name = os.path.basename(path)
    if name.endswith(".hip"):
        return [path]
    seen, work, out = {name}, [name], set()
    while work:
        h = work.pop()
        for f in glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.hpp")):
            if re.search(rf'#include\s+"{re.escape(h)}"', open(f).read()):
                b = os.path.basename(f)
                if b.endswith(".hip"):
                    out.add(f)
                elif b not in seen:
                    seen.add(b)
                    work.append(b)
    return sorted(out)


def variant_name(key: str, value) -> str:
    vals = value if isinstance(value, (tuple, list)) else (value,)
    return "_".join(f"{m[len('CVLHIP_'):].lower()}_{v}" for m, v in zip(macros(key), vals))


def build_variants(space=SPACE, verbose=True) -> dict:
    """Compile every variant library; returns {variant: info}."""
    from concurrent.futures import ThreadPoolExecutor
    from . import build as B
    B.build(verbose=False)                     # the default objects the variants link with
    jobs = []
    for macro, (values, prim) in space.items():
        cur = key_value(macro)
        files = set()
        for m in macros(macro):
            src = current_value(m)[1]
            files |= set(includers(src) if src else glob.glob(os.path.join(CSRC, "*.hip")))
        files = sorted(files)
        for v in values:
            if norm(v) == norm(cur):
                continue
            jobs.append((macro, v, prim, files))

    def one(job):
        macro, v, prim, files = job
        name = variant_name(macro, v)
        d = os.path.join(TUNE_DIR, name)              # the library (it travels to the GPU host)
        od = os.path.join(B.OBJ, "tune", name)        # its objects (they stay here)
        os.makedirs(d, exist_ok=True)
        os.makedirs(od, exist_ok=True)
        objs = {os.path.basename(o): o for o in glob.glob(os.path.join(B.OBJ, "*.hip.o"))}
        for f in files:
            o = os.path.join(od, os.path.basename(f) + ".o")
            r = subprocess.run([B.HIPCC, *B.FLAGS, *B.extra_flags(), *defines(macro, v), "-I", B.OBJ, "-c", f,
                                "-o", o], capture_output=True, text=True)
            if r.returncode != 0:
                return name, {"macro": macro, "value": v, "primitive": prim, "error": r.stderr[-800:]}
            objs[os.path.basename(f) + ".o"] = o
        lib = os.path.join(d, "libcvl_hip.so")
        r = subprocess.run([B.HIPCC, f"--offload-arch={B.ARCH}", "-shared", "-fPIC", "-o", lib, *objs.values()],
                           capture_output=True, text=True)
        if r.returncode != 0:
            return name, {"macro": macro, "value": v, "primitive": prim, "error": r.stderr[-800:]}
        return name, {"macro": macro, "value": v, "primitive": prim, "lib": os.path.relpath(lib, ROOT)}
    with ThreadPoolExecutor(max_workers=8) as ex:
        res = dict(ex.map(one, jobs))
    res["default"] = {"macro": None, "value": None, "primitive": None,
                      "lib": os.path.relpath(B.LIB, ROOT),
                      "defaults": {m: key_value(m) for m in space}}
    with open(os.path.join(TUNE_DIR, "variants.json"), "w") as f:
        json.dump(res, f, indent=1)
    if verbose:
        for k, v in res.items():
            print(k, "error" if "error" in v else "ok", flush=True)
    return res


# ------------------------------------------------------------------ GPU: one variant
def _time(fn, reps, torch):
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ts = []
    for r in range(reps + 2):
        torch.cuda.synchronize()
        e0.record(st)
        fn()
        e1.record(st)
        e1.synchronize()
        if r >= 2:
            ts.append(e0.elapsed_time(e1))
    ts.sort()
    return ts[len(ts) // 2]


def measure(primitive: str, reps: int = 7) -> dict:
    """Check `primitive` of the library this process loaded against the
    oracle on a small input, then time it at the benchmark's size."""
    import numpy as np
    import torch
    from neslgpu_amd.cvl import Cvl
    from oracle import oracle as O
    from . import dispatch as D
    cv = Cvl()
    L = cv.lib
    dev = cv.device
    rng = np.random.default_rng(3)

    def seg_case(n):
        lens = D.lengths("zipf", n, 1.5, rng, lmax=1 << 20)
        return lens, cv.segd(lens)

    ok = True
    # correctness, small (several tiles of every shape searched)
    n = 3_000_017
    if primitive in ("add_suz", "max_ruz", "pk2_luz", "smp_puz", "rku_luz"):
        x = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
        dx = torch.from_numpy(x).to(dev)
        if primitive == "add_suz":
            d = torch.empty_like(dx)
            L.add_suz(d, dx, n, cv.scratch(L.add_suz_scratch(n)))
            ok = np.array_equal(d.cpu().numpy(), O.scan("add", "z", x))
        elif primitive == "max_ruz":
            ok = L.max_ruz(dx, n, cv.scratch(L.max_ruz_scratch(n))) == int(x.max())
        elif primitive == "pk2_luz":
            f = (rng.random(n) < 0.5).astype(np.int32)
            k = int(f.sum())
            d = torch.empty(k, dtype=torch.int32, device=dev)
            L.pk2_luz(d, dx, torch.from_numpy(f).to(dev), n, k, cv.scratch(L.pk2_luz_scratch(n)))
            ok = np.array_equal(d.cpu().numpy(), O.pk2("z", x, f))
        elif primitive == "smp_puz":
            # one partition pass below 2^21 elements, two above; int and double
            for k in (100_003, n):
                p = rng.permutation(k).astype(np.int32)
                dp = torch.from_numpy(p).to(dev)
                d = torch.empty_like(dx[:k])
                L.smp_puz(d, dx[:k], dp, k, cv.scratch(L.smp_puz_scratch(k)))
                ok = ok and np.array_equal(d.cpu().numpy(), O.smp_pu("z", x[:k], p))
                y = rng.uniform(-1, 1, k)
                dy = torch.from_numpy(y).to(dev)
                e = torch.empty_like(dy)
                L.smp_pud(e, dy, dp, k, cv.scratch(L.smp_pud_scratch(k)))
                ok = ok and np.array_equal(e.cpu().numpy(), O.smp_pu("d", y, p))
        else:
            # int up, double down with ties, segmented int
            d = torch.empty_like(dx)
            L.rku_luz(d, dx, n, cv.scratch(L.rku_luz_scratch(n)))
            ok = np.array_equal(d.cpu().numpy(), O.rank(x, True))
            y = np.round(rng.uniform(-1e3, 1e3, n), 1)
            L.rkd_lud(d, torch.from_numpy(y).to(dev), n, cv.scratch(L.rkd_lud_scratch(n)))
            ok = ok and np.array_equal(d.cpu().numpy(), O.rank(y, False))
            lens, sd = seg_case(n)
            L.rku_lez(d, dx, sd, n, lens.size, cv.scratch(L.rku_lez_scratch(n, lens.size)))
            ok = ok and np.array_equal(d.cpu().numpy(), O.rank(x, True, lens))
    elif primitive == "add_sez":
        # the dispatched schedule and each flat arm, on two segment shapes
        for lens in (D.lengths("geometric", n, 256, rng), D.lengths("zipf", n, 1.5, rng, lmax=1 << 20)):
            sd, m = cv.segd(lens), lens.size
            x = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
            dx = torch.from_numpy(x).to(dev)
            want = O.segscan("add", "z", x, lens)
            d = torch.empty_like(dx)
            L.add_sez(d, dx, sd, n, m, cv.scratch(L.add_sez_scratch(n, m)))
            ok = ok and np.array_equal(d.cpu().numpy(), want)
            scr = cv.scratch(L.cvlhip_seg_scratch(n, m))
            for arm in range(3):
                L.cvlhip_seg_scan(D.OPS["add"], 0, arm, d, dx, sd, n, m, scr)
                ok = ok and np.array_equal(d.cpu().numpy(), want)
    else:
        lens, sd = seg_case(n)
        m = lens.size
        x = rng.uniform(-1, 1, n)
        dx = torch.from_numpy(x).to(dev)
        r = torch.empty(m, dtype=torch.float64, device=dev)
        if primitive == "max_red":
            L.max_red(r, dx, sd, n, m, cv.scratch(L.max_red_scratch(n, m)))
        elif primitive == "add_sed":
            # a float sum: the scan's order of additions is not the serial one
            sc = torch.empty_like(dx)
            L.add_sed(sc, dx, sd, n, m, cv.scratch(L.add_sed_scratch(n, m)))
            want = O.segscan("add", "d", x, lens)
            ok = np.allclose(sc.cpu().numpy(), want, rtol=0, atol=1e-9 * (1 + np.abs(want).max()))
        else:
            sc = torch.empty_like(dx)
            L.cvlhip_add_sed_max_red(sc, r, dx, sd, n, m, cv.scratch(L.cvlhip_add_sed_max_red_scratch(n, m)))
        if primitive != "add_sed":
            ok = np.array_equal(r.cpu().numpy().view(np.int64), O.segreduce("max", "d", x, lens).view(np.int64))
    ok = bool(ok) and L.cvlhip_check_device_errors() == 0
    # timing at the benchmark's size
    if primitive in ("max_red", "add_sed", "add_sed_max_red"):
        n = 1 << 26
        x = torch.empty(n, dtype=torch.float64, device=dev)
        L.cvlhip_fill_splitmix_f64(x, n, 0x4E45534C00000003, -1.0, 1.0)
        lens = D.lengths("zipf", n, 1.5, np.random.default_rng(0x4E45534C00000003 & 0x7FFFFFFF), lmax=1 << 20)
        sd, m = cv.segd(lens), lens.size
        r = torch.empty(m, dtype=torch.float64, device=dev)
        if primitive == "max_red":
            scr = cv.scratch(L.max_red_scratch(n, m))
            fn = lambda: L.max_red(r, x, sd, n, m, scr)
        elif primitive == "add_sed":
            sc = torch.empty_like(x)
            scr = cv.scratch(L.add_sed_scratch(n, m))
            fn = lambda: L.add_sed(sc, x, sd, n, m, scr)
        else:
            sc = torch.empty_like(x)
            scr = cv.scratch(L.cvlhip_add_sed_max_red_scratch(n, m))
            fn = lambda: L.cvlhip_add_sed_max_red(sc, r, x, sd, n, m, scr)
    else:
        n = (1 << 26) if primitive == "rku_luz" else (1 << 28)
        s = torch.empty(n, dtype=torch.int32, device=dev)
        L.cvlhip_fill_splitmix_u32(s, n, 0x4E45534C00000002)
        d = torch.empty_like(s)
        if primitive == "add_suz":
            scr = cv.scratch(L.add_suz_scratch(n))
            fn = lambda: L.add_suz(d, s, n, scr)
        elif primitive == "add_sez":
            lens = D.lengths("geometric", n, 256, np.random.default_rng(4))
            sd, m = cv.segd(lens), lens.size
            scr = cv.scratch(L.add_sez_scratch(n, m))
            fn = lambda: L.add_sez(d, s, sd, n, m, scr)
        elif primitive == "max_ruz":
            scr = cv.scratch(L.max_ruz_scratch(n))
            fn = lambda: L.max_ruz(s, n, scr)
        elif primitive == "pk2_luz":
            f = (s < 0).to(torch.int32)
            k = int(f.sum().item())
            scr = cv.scratch(L.pk2_luz_scratch(n))
            fn = lambda: L.pk2_luz(d, s, f, n, k, scr)
        elif primitive == "smp_puz":
            g = torch.Generator(device=dev)
            g.manual_seed(5)
            p = torch.randperm(n, device=dev, generator=g, dtype=torch.int64).to(torch.int32)
            scr = cv.scratch(L.smp_puz_scratch(n))
            fn = lambda: L.smp_puz(d, s, p, n, scr)
        else:
            scr = cv.scratch(L.rku_luz_scratch(n))
            fn = lambda: L.rku_luz(d, s, n, scr)
    return {"primitive": primitive, "correct": ok, "ms": round(_time(fn, reps, torch), 4), "n": n}


def run_variants(out_path: str, reps: int = 7) -> dict:
    """Every built variant (and the default) in a process of its own."""
    with open(os.path.join(TUNE_DIR, "variants.json")) as f:
        variants = json.load(f)
    prims = sorted({v["primitive"] for v in variants.values() if v.get("primitive")})
    results = {}
    todo = [("default", p) for p in prims] + [(k, v["primitive"]) for k, v in variants.items()
                                            if k != "default" and "lib" in v]
    for name, prim in todo:
        lib = os.path.join(ROOT, variants[name]["lib"])
        env = dict(os.environ, CVLHIP_LIBRARY=lib)
        t0 = time.time()
        r = subprocess.run([sys.executable, "-m", "neslgpu_amd.tune", "_measure", prim, str(reps)], cwd=ROOT,
                           env=env, capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            rec = {"primitive": prim, "error": r.stderr[-600:], "rc": r.returncode}
        else:
            rec = json.loads(r.stdout.strip().splitlines()[-1])
        rec["seconds"] = round(time.time() - t0, 1)
        results[f"{name}:{prim}" if name == "default" else name] = dict(rec, **{k: variants[name].get(k)
                                                                              for k in ("macro", "value")})
        print(json.dumps({name: results[f"{name}:{prim}" if name == "default" else name]}), flush=True)
        if r.returncode not in (0, 1):
            break                      # a crash of the GPU process: stop here
    out = {"variants": variants, "results": results, "reps": reps}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    return out


# ------------------------------------------------------------------ pick
def pick(path: str) -> dict:
    with open(path) as f:
        rec = json.load(f)
    res, variants = rec["results"], rec["variants"]
    defaults = variants["default"]["defaults"]
    chosen, rows = {}, []
    for macro, (values, prim) in SPACE.items():
        if macro not in defaults:               # not searched in this run: keeps its value
            chosen[macro] = norm(key_value(macro))
            continue
        base = res.get(f"default:{prim}", {})
        best_v, best_ms = norm(defaults[macro]), base.get("ms") if base.get("correct") else None
        rows.append((macro, best_v, prim, base.get("ms"), base.get("correct"), "current"))
        for v in values:
            if norm(v) == norm(defaults[macro]):
                continue
            r = res.get(variant_name(macro, v))
            if not r:
                rows.append((macro, v, prim, None, None, "not built"))
                continue
            rows.append((macro, v, prim, r.get("ms"), r.get("correct"), r.get("error", "")[:60]))
            # a change must beat the current value by more than the run-to-run spread
            if r.get("correct") and r.get("ms") and best_ms and r["ms"] < 0.98 * best_ms:
                best_v, best_ms = v, r["ms"]
        chosen[macro] = best_v
    lines = ["// tuned.hpp -- tile and workgroup shapes chosen by the autotuner",
             "// (python -m neslgpu_amd.tune pick), force-included into every HIP",
             "// source by neslgpu_amd/build.py.  Each value is the fastest correct",
             "// one of its macro's candidates on MI355X, the others at their values.",
             "#pragma once"]
    for key, value in chosen.items():           # a -D on the command line still wins
        vals = value if isinstance(value, tuple) else (value,)
        for m, v in zip(macros(key), vals):
            lines += [f"#ifndef {m}", f"#define {m} {v}", "#endif"]
    with open(TUNED, "w") as f:
        f.write("\n".join(lines) + "\n")
    md = ["# Tile-shape autotuning (MI355X)", "",
          f"Command: `python -m neslgpu_amd.tune build && python -m neslgpu_amd.tune run {os.path.basename(path)}"
          f" && python -m neslgpu_amd.tune pick {os.path.basename(path)}`", "",
          "Coordinate search: each candidate of one macro with every other macro at its value; median of "
          f"{rec.get('reps')} timed calls (HIP events) at the benchmark's size, after an oracle check on "
          "3,000,017 elements.  A candidate replaces the current value only when it is correct and more than "
          "2% faster.", "",
          "| macro | value | judged by | ms | correct | note |", "|---|---:|---|---:|---|---|"]
    for macro, v, prim, ms, ok, note in rows:
        md.append(f"| {macro} | {v} | {prim} | {ms if ms is not None else ''} | {ok if ok is not None else ''} | "
                  f"{note} |")
    md += ["", "Chosen: " + ", ".join(f"{m} = {v}" for m, v in chosen.items())]
    with open(RECORD, "w") as f:
        f.write("\n".join(md) + "\n")
    return chosen


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    cmd = argv[0] if argv else "build"
    if cmd == "build":
        # optional: the SPACE keys to search (all by default)
        build_variants({k: SPACE[k] for k in argv[1:]} if len(argv) > 1 else SPACE)
    elif cmd == "run":
        run_variants(argv[1] if len(argv) > 1 else os.path.join(TUNE_DIR, "results.json"))
    elif cmd == "pick":
        print(json.dumps(pick(argv[1])))
    elif cmd == "_measure":
        sys.path.insert(0, ROOT)
        print(json.dumps(measure(argv[1], int(argv[2]))), flush=True)
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main()
