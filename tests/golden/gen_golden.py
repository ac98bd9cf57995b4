"""Writes the golden vectors under tests/golden/ from the reference serial CVL.

Run in a container where the reference sources exist:
    make -C oracle && python tests/golden/gen_golden.py

It loads oracle/_ref/libcvl_serial.so (built by oracle/Makefile from
nesl/cvl/serial/*.c), calls each reference function on seeded inputs through
its own C API, and saves inputs and outputs as .npz fixtures (plain arrays,
no pickles).  Fixture keys are "<case>/<field>".  The GPU box never runs
this script; tests read only the .npz files it wrote.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from oracle import refcvl  # noqa: E402

RNG = np.random.default_rng(0x4E45534C)
SIZES = [0, 1, 2, 7, 64, 65, 517, 1031]

Z_OPS = ["add", "mul", "min", "max", "and", "ior", "xor"]
D_OPS = ["add", "mul", "min", "max"]
B_OPS = ["and", "ior", "xor"]


def rand_z(n, lo=-(2 ** 31), hi=2 ** 31):
    return RNG.integers(lo, hi, size=n, dtype=np.int64).astype(np.int32)


def rand_d(n):
    return RNG.uniform(-4.0, 4.0, size=n)


def rand_b(n, p=0.5):
    return (RNG.random(n) < p).astype(np.int32)


def rand_of(ty, n):
    return {"z": rand_z, "d": rand_d, "b": rand_b}[ty](n)


def rand_lens(n, kind):
    """Segment lengths summing to n (some empty segments included)."""
    if kind == "one":
        return np.array([n], np.int32)
    out = []
    left = n
    while left > 0:
        L = int(RNG.geometric(1.0 / (4 if kind == "short" else 40))) - 1
        L = min(L, left)
        out.append(L)
        left -= L
    if kind == "short":
        out += [0, 0]
        RNG.shuffle(out)
    return np.array(out if out else [0], np.int32)


def seg_cases():
    yield "fixed", np.array([3, 0, 5, 1, 0, 0, 7], np.int32)
    for n in (64, 1031):
        for kind in ("short", "long", "one"):
            yield f"{kind}{n}", rand_lens(n, kind)
    yield "empty", np.zeros(0, np.int32)


def scratch(units):
    return refcvl.scratch_for(units)


def dt(ty):
    return np.float64 if ty == "d" else np.int32


def gen_scans(out):
    for ty, ops in (("z", Z_OPS), ("d", D_OPS), ("b", B_OPS)):
        for op in ops:
            for n in SIZES:
                s = rand_of(ty, n)
                if op == "mul" and ty == "d":
                    s = RNG.uniform(0.5, 1.5, size=n)
                d = np.zeros(n, dt(ty))
                refcvl.call(f"{op}_su{ty}", d, s, n, scratch(0))
                out[f"{op}_su{ty}/n{n}/s"] = s
                out[f"{op}_su{ty}/n{n}/d"] = d
                r = refcvl.call(f"{op}_ru{ty}", s, n, scratch(0))
                out[f"{op}_ru{ty}/n{n}/s"] = s
                out[f"{op}_ru{ty}/n{n}/r"] = np.array([r], dt(ty))
            for name, lens in seg_cases():
                n = int(lens.sum())
                m = lens.size
                s = rand_of(ty, n)
                if op == "mul" and ty == "d":
                    s = RNG.uniform(0.5, 1.5, size=n)
                d = np.zeros(n, dt(ty))
                refcvl.call(f"{op}_se{ty}", d, s, lens, n, m, scratch(0))
                out[f"{op}_se{ty}/{name}/s"] = s
                out[f"{op}_se{ty}/{name}/lens"] = lens
                out[f"{op}_se{ty}/{name}/d"] = d
                r = np.zeros(m, dt(ty))
                refcvl.call(f"{op}_re{ty}", r, s, lens, n, m, scratch(0))
                out[f"{op}_re{ty}/{name}/s"] = s
                out[f"{op}_re{ty}/{name}/lens"] = lens
                out[f"{op}_re{ty}/{name}/d"] = r


def gen_pack(out):
    for n in SIZES:
        f = rand_b(n, 0.5)
        out[f"pk1_luv/n{n}/f"] = f
        out[f"pk1_luv/n{n}/r"] = np.array([refcvl.call("pk1_luv", f, n, scratch(0))], np.int32)
        for ty in "zbd":
            s = rand_of(ty, n)
            cnt = int(f.sum())
            d = np.zeros(cnt, dt(ty))
            refcvl.call(f"pk2_lu{ty}", d, s, f, n, cnt, scratch(0))
            out[f"pk2_lu{ty}/n{n}/s"] = s
            out[f"pk2_lu{ty}/n{n}/f"] = f
            out[f"pk2_lu{ty}/n{n}/d"] = d
    for name, lens in seg_cases():
        n = int(lens.sum()); m = lens.size
        f = rand_b(n, 0.3)
        ds = np.zeros(m, np.int32)
        refcvl.call("pk1_lev", ds, f, lens, n, m, scratch(0))
        out[f"pk1_lev/{name}/f"] = f
        out[f"pk1_lev/{name}/lens"] = lens
        out[f"pk1_lev/{name}/d"] = ds
        for ty in "zbd":
            s = rand_of(ty, n)
            nd = int(ds.sum())
            d = np.zeros(nd, dt(ty))
            refcvl.call(f"pk2_le{ty}", d, s, f, lens, n, m, ds, nd, m, scratch(0))
            out[f"pk2_le{ty}/{name}/s"] = s
            out[f"pk2_le{ty}/{name}/f"] = f
            out[f"pk2_le{ty}/{name}/lens"] = lens
            out[f"pk2_le{ty}/{name}/d"] = d


def seg_perm(lens):
    """Per-segment random permutations (index relative to the segment)."""
    parts = [RNG.permutation(int(L)).astype(np.int32) for L in lens]
    return np.concatenate(parts) if parts else np.zeros(0, np.int32)


def gen_permute(out):
    for n in SIZES:
        idx = RNG.permutation(n).astype(np.int32)
        f = rand_b(n, 0.5)
        for ty in "zbd":
            s = rand_of(ty, n)
            d = np.zeros(n, dt(ty))
            refcvl.call(f"smp_pu{ty}", d, s, idx, n, scratch(0))
            out[f"smp_pu{ty}/n{n}/s"] = s; out[f"smp_pu{ty}/n{n}/i"] = idx
            out[f"smp_pu{ty}/n{n}/d"] = d
            # back-permute from a longer source with repeated indices
            ns = 2 * n + 3
            src = rand_of(ty, ns)
            gi = RNG.integers(0, ns, size=n).astype(np.int32)
            d = np.zeros(n, dt(ty))
            refcvl.call(f"bck_pu{ty}", d, src, gi, ns, n, scratch(0))
            out[f"bck_pu{ty}/n{n}/s"] = src; out[f"bck_pu{ty}/n{n}/i"] = gi
            out[f"bck_pu{ty}/n{n}/d"] = d
            # flag permute / flag back-permute / default permutes
            d = np.zeros(n, dt(ty))
            refcvl.call(f"fpm_pu{ty}", d, s, idx, f, n, n, scratch(0))
            out[f"fpm_pu{ty}/n{n}/s"] = s; out[f"fpm_pu{ty}/n{n}/i"] = idx
            out[f"fpm_pu{ty}/n{n}/f"] = f; out[f"fpm_pu{ty}/n{n}/d"] = d
            d = np.full(n, 7, dt(ty))
            refcvl.call(f"bfp_pu{ty}", d, src, gi, f, ns, n, scratch(0))
            out[f"bfp_pu{ty}/n{n}/s"] = src; out[f"bfp_pu{ty}/n{n}/i"] = gi
            out[f"bfp_pu{ty}/n{n}/f"] = f; out[f"bfp_pu{ty}/n{n}/d"] = d
            v = rand_of(ty, ns)
            sub = RNG.permutation(ns)[:n].astype(np.int32)
            d = np.zeros(ns, dt(ty))
            refcvl.call(f"dpe_pu{ty}", d, s, sub, v, n, ns, scratch(0))
            out[f"dpe_pu{ty}/n{n}/s"] = s; out[f"dpe_pu{ty}/n{n}/i"] = sub
            out[f"dpe_pu{ty}/n{n}/v"] = v; out[f"dpe_pu{ty}/n{n}/d"] = d
            d = np.zeros(ns, dt(ty))
            refcvl.call(f"dfp_pu{ty}", d, s, sub, f, v, n, ns, scratch(0))
            out[f"dfp_pu{ty}/n{n}/s"] = s; out[f"dfp_pu{ty}/n{n}/i"] = sub
            out[f"dfp_pu{ty}/n{n}/f"] = f; out[f"dfp_pu{ty}/n{n}/v"] = v
            out[f"dfp_pu{ty}/n{n}/d"] = d
    for name, lens in seg_cases():
        n = int(lens.sum()); m = lens.size
        idx = seg_perm(lens)
        f = rand_b(n, 0.5)
        for ty in "zbd":
            s = rand_of(ty, n)
            d = np.zeros(n, dt(ty))
            refcvl.call(f"smp_pe{ty}", d, s, idx, lens, n, m, scratch(0))
            k = f"smp_pe{ty}/{name}"
            out[k + "/s"] = s; out[k + "/i"] = idx; out[k + "/lens"] = lens; out[k + "/d"] = d
            # segmented back-permute: source segments twice as long
            lens_s = (2 * lens + 1).astype(np.int32)
            ns = int(lens_s.sum())
            src = rand_of(ty, ns)
            gi = np.concatenate([RNG.integers(0, int(Ls), size=int(L)) for L, Ls in
                                 zip(lens, lens_s)]).astype(np.int32) if m else np.zeros(0, np.int32)
            d = np.zeros(n, dt(ty))
            refcvl.call(f"bck_pe{ty}", d, src, gi, lens_s, ns, m, lens, n, m, scratch(0))
            k = f"bck_pe{ty}/{name}"
            out[k + "/s"] = src; out[k + "/i"] = gi; out[k + "/lens_s"] = lens_s
            out[k + "/lens_d"] = lens; out[k + "/d"] = d
            # flagged segmented permutes into destination segments lens_s
            sub = np.concatenate([RNG.permutation(int(Ls))[:int(L)] for L, Ls in
                                  zip(lens, lens_s)]).astype(np.int32) if m else np.zeros(0, np.int32)
            d = np.zeros(ns, dt(ty))
            refcvl.call(f"fpm_pe{ty}", d, s, sub, f, lens, n, m, lens_s, ns, m, scratch(0))
            k = f"fpm_pe{ty}/{name}"
            out[k + "/s"] = s; out[k + "/i"] = sub; out[k + "/f"] = f
            out[k + "/lens_s"] = lens; out[k + "/lens_d"] = lens_s; out[k + "/d"] = d
            v = rand_of(ty, ns)
            d = np.zeros(ns, dt(ty))
            refcvl.call(f"dpe_pe{ty}", d, s, sub, v, lens, n, m, lens_s, ns, m, scratch(0))
            k = f"dpe_pe{ty}/{name}"
            out[k + "/s"] = s; out[k + "/i"] = sub; out[k + "/v"] = v
            out[k + "/lens_s"] = lens; out[k + "/lens_d"] = lens_s; out[k + "/d"] = d
            d = np.zeros(ns, dt(ty))
            refcvl.call(f"dfp_pe{ty}", d, s, sub, f, v, lens, n, m, lens_s, ns, m, scratch(0))
            k = f"dfp_pe{ty}/{name}"
            out[k + "/s"] = s; out[k + "/i"] = sub; out[k + "/f"] = f; out[k + "/v"] = v
            out[k + "/lens_s"] = lens; out[k + "/lens_d"] = lens_s; out[k + "/d"] = d
            # flagged segmented back-permute: destination lens, source lens_s
            d = np.zeros(n, dt(ty))
            refcvl.call(f"bfp_pe{ty}", d, src, gi, f, lens_s, ns, m, lens, n, m, scratch(0))
            k = f"bfp_pe{ty}/{name}"
            out[k + "/s"] = src; out[k + "/i"] = gi; out[k + "/f"] = f
            out[k + "/lens_s"] = lens_s; out[k + "/lens_d"] = lens; out[k + "/d"] = d


def special_doubles(n):
    base = np.array([0.0, -0.0, 1.0, -1.0, 5e-324, -5e-324, 1e308, -1e308,
                     np.inf, -np.inf, 2.5, -2.5, 1.0, 0.0])
    return np.concatenate([base, RNG.uniform(-3, 3, n)])


def gen_rank(out):
    for n in SIZES:
        for ty in "zd":
            if ty == "z":
                s = rand_z(n, -50, 50) if n % 2 else rand_z(n)
            else:
                s = special_doubles(n) if n > 1 else rand_d(n)
            nn = s.size
            for up in ("u", "d"):
                d = np.zeros(nn, np.int32)
                refcvl.call(f"rk{up}_lu{ty}", d, s, nn, scratch(2 * nn + 1))
                k = f"rk{up}_lu{ty}/n{n}"
                out[k + "/s"] = s; out[k + "/d"] = d
    for name, lens in seg_cases():
        n = int(lens.sum()); m = lens.size
        for ty in "zd":
            s = rand_z(n, -20, 20) if ty == "z" else np.round(rand_d(n), 1)
            for up in ("u", "d"):
                d = np.zeros(n, np.int32)
                refcvl.call(f"rk{up}_le{ty}", d, s, lens, n, m, scratch(2 * n + 1))
                k = f"rk{up}_le{ty}/{name}"
                out[k + "/s"] = s; out[k + "/lens"] = lens; out[k + "/d"] = d


def gen_misc(out):
    for n in SIZES:
        d = np.zeros(n, np.int32)
        init, stride = int(rand_z(1)[0]), int(rand_z(1, -1000, 1000)[0])
        refcvl.call("ind_luz", d, init, stride, n, scratch(0))
        out[f"ind_luz/n{n}/args"] = np.array([init, stride], np.int32)
        out[f"ind_luz/n{n}/d"] = d
    for name, lens in seg_cases():
        n = int(lens.sum()); m = lens.size
        init = rand_z(m); stride = rand_z(m, -100, 100)
        d = np.zeros(n, np.int32)
        refcvl.call("ind_lez", d, init, stride, lens, n, m, scratch(0))
        k = f"ind_lez/{name}"
        out[k + "/init"] = init; out[k + "/stride"] = stride; out[k + "/lens"] = lens
        out[k + "/d"] = d
        for ty in "zbd":
            v = rand_of(ty, m)
            d = np.zeros(n, dt(ty))
            refcvl.call(f"dis_ve{ty}", d, v, lens, n, m, scratch(0))
            k = f"dis_ve{ty}/{name}"
            out[k + "/v"] = v; out[k + "/lens"] = lens; out[k + "/d"] = d
            nz = lens > 0
            lens_nz = lens[nz]
            mn = lens_nz.size
            nn = int(lens_nz.sum())
            s = rand_of(ty, nn)
            ii = np.array([RNG.integers(0, L) for L in lens_nz], np.int32)
            d = np.zeros(mn, dt(ty))
            refcvl.call(f"ext_ve{ty}", d, s, ii, lens_nz, nn, mn, scratch(0))
            k = f"ext_ve{ty}/{name}"
            out[k + "/s"] = s; out[k + "/i"] = ii; out[k + "/lens"] = lens_nz; out[k + "/d"] = d
            d = s.copy()
            vv = rand_of(ty, mn)
            refcvl.call(f"rep_ve{ty}", d, ii, vv, lens_nz, nn, mn, scratch(0))
            k = f"rep_ve{ty}/{name}"
            out[k + "/s"] = s; out[k + "/i"] = ii; out[k + "/v"] = vv
            out[k + "/lens"] = lens_nz; out[k + "/d"] = d


EW2_Z = ["max", "min", "add", "sub", "mul", "div", "grt", "les", "geq", "leq", "eql",
         "neq", "lsh", "rsh", "mod", "and", "ior", "xor"]
EW2_D = ["max", "min", "add", "sub", "mul", "div", "grt", "les", "geq", "leq", "eql", "neq"]
EW2_B = ["eql", "neq", "and", "ior", "xor"]
EW1 = {"not_wub": "b", "not_wuz": "z", "int_wud": "d", "int_wub": "b", "dbl_wuz": "z",
       "boo_wuz": "z", "flr_wud": "d", "cei_wud": "d", "trn_wud": "d", "rou_wud": "d",
       "exp_wud": "d", "log_wud": "dpos", "sqt_wud": "dpos", "sin_wud": "d", "cos_wud": "d",
       "tan_wud": "d", "asn_wud": "dunit", "acs_wud": "dunit", "atn_wud": "d",
       "snh_wud": "d", "csh_wud": "d", "tnh_wud": "d"}
EW1_OUT = {"dbl_wuz": "d", "exp_wud": "d", "log_wud": "d", "sqt_wud": "d", "sin_wud": "d",
           "cos_wud": "d", "tan_wud": "d", "asn_wud": "d", "acs_wud": "d", "atn_wud": "d",
           "snh_wud": "d", "csh_wud": "d", "tnh_wud": "d"}


def gen_elwise(out):
    n = 1000
    for op in EW2_Z:
        a = rand_z(n)
        b = rand_z(n)
        if op in ("div", "mod"):
            b = rand_z(n, 1, 1000) * np.where(RNG.random(n) < 0.5, 1, -1).astype(np.int32)
            a[a == np.iinfo(np.int32).min] = 0
        if op in ("lsh", "rsh"):
            b = rand_z(n, 0, 32)
        if op in ("eql", "neq"):
            b = np.where(RNG.random(n) < 0.5, a, b).astype(np.int32)
        d = np.zeros(n, np.int32)
        name = f"{op}_wuz"
        refcvl.call(name, d, a, b, n, scratch(0))
        out[f"{name}/a"] = a; out[f"{name}/b"] = b; out[f"{name}/d"] = d
    for op in EW2_D:
        a = rand_d(n); b = rand_d(n)
        if op in ("eql", "neq", "geq", "leq"):
            b = np.where(RNG.random(n) < 0.5, a, b)
        cmp = op in ("grt", "les", "geq", "leq", "eql", "neq")
        d = np.zeros(n, np.int32 if cmp else np.float64)
        name = f"{op}_wud"
        refcvl.call(name, d, a, b, n, scratch(0))
        out[f"{name}/a"] = a; out[f"{name}/b"] = b; out[f"{name}/d"] = d
    for op in EW2_B:
        a = rand_b(n); b = rand_b(n)
        d = np.zeros(n, np.int32)
        name = f"{op}_wub"
        refcvl.call(name, d, a, b, n, scratch(0))
        out[f"{name}/a"] = a; out[f"{name}/b"] = b; out[f"{name}/d"] = d
    for name, kind in EW1.items():
        if kind == "b":
            s = rand_b(n)
        elif kind == "z":
            s = rand_z(n)
        elif kind == "dpos":
            s = RNG.uniform(1e-3, 100.0, n)
        elif kind == "dunit":
            s = RNG.uniform(-1.0, 1.0, n)
        else:
            s = RNG.uniform(-20.0, 20.0, n)
            if name in ("flr_wud", "cei_wud", "trn_wud", "rou_wud"):
                s[:8] = [-2.5, -1.5, -0.5, 0.5, 1.5, 2.5, -0.49999999999999994, 3.0]
        d = np.zeros(n, np.float64 if EW1_OUT.get(name) == "d" else np.int32)
        refcvl.call(name, d, s, n, scratch(0))
        out[f"{name}/s"] = s; out[f"{name}/d"] = d
    for ty in "zbd":
        f = rand_b(n); a = rand_of(ty, n); b = rand_of(ty, n)
        d = np.zeros(n, dt(ty))
        refcvl.call(f"sel_wu{ty}", d, f, a, b, n, scratch(0))
        out[f"sel_wu{ty}/f"] = f; out[f"sel_wu{ty}/a"] = a; out[f"sel_wu{ty}/b"] = b
        out[f"sel_wu{ty}/d"] = d


def gen_rand(out):
    # First RAND of an unseeded process: the reference never calls srand48 on
    # Linux (its rnd_foz calls srandom), so lrand48 runs from glibc's initial
    # state.  Nothing else in this process has called lrand48.
    s = rand_z(4099, 1, 2 ** 31 - 1)
    s[:5] = [1, 2, 10, 1000, 2147483647]
    d = np.zeros(s.size, np.int32)
    refcvl.call("rnd_wuz", d, s, s.size, scratch(0))
    out["rnd_wuz/first/s"] = s; out["rnd_wuz/first/d"] = d
    s2 = rand_z(1000, 1, 100)
    d2 = np.zeros(s2.size, np.int32)
    refcvl.call("rnd_wuz", d2, s2, s2.size, scratch(0))
    out["rnd_wuz/second/s"] = s2; out["rnd_wuz/second/d"] = d2


def gen_rank_asis(out):
    """The as-is x86-64 build of the reference rank.c (no word-order define)."""
    if not refcvl.available(refcvl.ASIS_PATH):
        return
    s = np.array([10.0, -5.0, 0.0, 10.0, -10.0, 100.0, 10.0])
    d = np.zeros(7, np.int32)
    refcvl.call("rku_lud", d, s, 7, scratch(16), path=refcvl.ASIS_PATH)
    out["asis_rku_lud/kat/s"] = s; out["asis_rku_lud/kat/d"] = d


def main():
    if not refcvl.available():
        sys.exit("oracle/_ref/libcvl_serial.so missing: run `make -C oracle` first")
    fams = {"scans": gen_scans, "pack": gen_pack, "permute": gen_permute,
            "rank": gen_rank, "misc": gen_misc, "elwise": gen_elwise}
    for fam, fn in fams.items():
        out: dict[str, np.ndarray] = {}
        fn(out)
        np.savez_compressed(os.path.join(HERE, f"{fam}.npz"), **out)
        print(fam, len(out), "arrays")
    out = {}
    gen_rand(out)
    gen_rank_asis(out)
    np.savez_compressed(os.path.join(HERE, "rand.npz"), **out)
    print("rand", len(out), "arrays")


if __name__ == "__main__":
    main()
