"""GPU parity against the golden vectors of the reference serial CVL.

Every call goes through the C-ABI (libcvl_hip.so) with the reference
argument order.  Integer, boolean, index, pack, permute, rank and double
min/max results must be bit-exact; double +-scans and reductions are
checked against the compensated (Neumaier) oracle with the tolerance
2(ceil(log2 L)+2) 2^-53 sum|x|, double *-scans against the serial result with
relative tolerance 4 L 2^-53; libm transcendentals within 4 ulp.
"""
import sys

import numpy as np
import pytest

from conftest import golden_cases, load_golden
from gpu_util import eq, fp_add_tol, host, np_dtype, seg_lengths_per_element
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cv():
    from neslgpu_amd.cvl import Cvl
    return Cvl()


def note(*a):
    print("case:", *a, flush=True)
    sys.stdout.flush()


def dev(cv, a, ty=None):
    import torch
    a = np.ascontiguousarray(a)
    if a.dtype == np.float64:
        return torch.from_numpy(a.copy()).to(cv.device)
    return torch.from_numpy(a.astype(np.int32)).to(cv.device)


def out(cv, n, ty):
    import torch
    return torch.zeros(max(int(n), 1), dtype=torch.float64 if ty == "d" else torch.int32, device=cv.device)


def check_float_add(got, s, lens, what):
    ref, absp = O.segscan_add_neumaier(s, lens)
    tol = np.array([fp_add_tol(a, L) for a, L in zip(absp, seg_lengths_per_element(lens))])
    err = np.abs(got - ref)
    assert np.all(err <= tol), (what, float(err.max()), float(tol[np.argmax(err - tol)]))


def check_float_add_reduce(got, s, lens, what):
    ref, absum = O.segreduce_add_neumaier(s, lens)
    tol = np.array([fp_add_tol(a, L) for a, L in zip(absum, lens)])
    err = np.abs(got - ref)
    assert np.all(err <= tol), (what, float(err.max()))


def check_float_mul(got, want, L, what):
    err = np.abs(got - want)
    tol = 4.0 * max(L, 2) * 2.0 ** -53 * np.abs(want) + 1e-300
    assert np.all(err <= tol), (what, float((err / np.maximum(np.abs(want), 1e-300)).max()))


def test_scans_reduces(cv):
    lib = cv.lib
    cases = golden_cases(load_golden("scans"), "")
    for (fn, case), f in sorted(cases.items()):
        op, kind, ty = fn[:3], fn[4:6], fn[6]
        s = f["s"]
        n = s.size
        note(fn, case, n)
        ds = dev(cv, s)
        if kind == "su":
            d = out(cv, n, ty)
            getattr(lib, fn)(d, ds, n, cv.scratch(getattr(lib, fn + "_scratch")(n)))
            got = host(d)[:n]
            if ty == "d" and op == "add":
                check_float_add(got, s, np.array([n]), fn + case)
            elif ty == "d" and op == "mul":
                check_float_mul(got, f["d"], n, fn + case)
            else:
                eq(got, f["d"], fn + case)
        elif kind == "ru":
            r = getattr(lib, fn)(ds, n, cv.scratch(getattr(lib, fn + "_scratch")(n)))
            want = f["r"][0]
            if ty == "d" and op == "add":
                check_float_add_reduce(np.array([r]), s, np.array([n]), fn + case)
            elif ty == "d" and op == "mul":
                check_float_mul(np.array([r]), np.array([want]), n, fn + case)
            else:
                eq(np.array([r], np_dtype(ty)), f["r"], fn + case)
        elif kind in ("se", "re"):
            lens = f["lens"]
            m = lens.size
            sd = cv.segd(lens)
            scr = cv.scratch(getattr(lib, fn + "_scratch")(n, m))
            if kind == "se":
                d = out(cv, n, ty)
                getattr(lib, fn)(d, ds, sd, n, m, scr)
                got = host(d)[:n]
                if ty == "d" and op == "add":
                    check_float_add(got, s, lens, fn + case)
                elif ty == "d" and op == "mul":
                    check_float_mul(got, f["d"], int(lens.max()) if m else 1, fn + case)
                else:
                    eq(got, f["d"], fn + case)
            else:
                d = out(cv, m, ty)
                getattr(lib, fn)(d, ds, sd, n, m, scr)
                got = host(d)[:m]
                if ty == "d" and op == "add":
                    check_float_add_reduce(got, s, lens, fn + case)
                elif ty == "d" and op == "mul":
                    check_float_mul(got, f["d"], int(lens.max()) if m else 1, fn + case)
                else:
                    eq(got, f["d"], fn + case)
    assert lib.cvlhip_check_device_errors() == 0


def test_pack(cv):
    lib = cv.lib
    cases = golden_cases(load_golden("pack"), "")
    for (fn, case), f in sorted(cases.items()):
        note(fn, case)
        if fn == "pk1_luv":
            n = f["f"].size
            r = lib.pk1_luv(dev(cv, f["f"]), n, cv.scratch(lib.pk1_luv_scratch(n)))
            assert r == int(f["r"][0]), (case, r)
        elif fn == "pk1_lev":
            lens = f["lens"]; n = int(lens.sum()); m = lens.size
            sd = cv.segd(lens)
            d = out(cv, m, "z")
            lib.pk1_lev(d, dev(cv, f["f"]), sd, n, m, cv.scratch(lib.pk1_lev_scratch(n, m)))
            eq(host(d)[:m], f["d"], fn + case)
        elif fn.startswith("pk2_lu"):
            ty = fn[6]
            n = f["s"].size
            cnt = f["d"].size
            d = out(cv, cnt, ty)
            getattr(lib, fn)(d, dev(cv, f["s"]), dev(cv, f["f"]), n, cnt,
                                cv.scratch(getattr(lib, fn + "_scratch")(n)))
            eq(host(d)[:cnt], f["d"], fn + case)
        else:
            ty = fn[6]
            lens = f["lens"]; n = int(lens.sum()); m = lens.size
            dlens = O.pk1_lev(f["f"], lens)
            nd = int(dlens.sum())
            sd = cv.segd(lens); sdd = cv.segd(dlens)
            d = out(cv, nd, ty)
            getattr(lib, fn)(d, dev(cv, f["s"]), dev(cv, f["f"]), sd, n, m, sdd, nd, m,
                                cv.scratch(getattr(lib, fn + "_scratch")(n, m, nd, m)))
            eq(host(d)[:nd], f["d"], fn + case)


def test_permutes(cv):
    lib = cv.lib
    cases = golden_cases(load_golden("permute"), "")
    for (fn, case), f in sorted(cases.items()):
        ty = fn[6]
        kind = fn[:6]
        want = f["d"]
        note(fn, case, want.size)
        s = dev(cv, f["s"]); i = dev(cv, f["i"])
        fl = dev(cv, f["f"]) if "f" in f else None
        v = dev(cv, f["v"]) if "v" in f else None
        nd = want.size
        ns = f["s"].size
        if kind.endswith("pu"):
            if kind == "smp_pu":
                d = out(cv, nd, ty)
                getattr(lib, fn)(d, s, i, ns, cv.scratch(0))
            elif kind == "bck_pu":
                d = out(cv, nd, ty)
                getattr(lib, fn)(d, s, i, ns, nd, cv.scratch(0))
            elif kind == "fpm_pu":
                d = out(cv, nd, ty)
                getattr(lib, fn)(d, s, i, fl, ns, nd, cv.scratch(getattr(lib, fn + "_scratch")(ns, nd)))
            elif kind == "bfp_pu":
                d = out(cv, nd, ty)
                getattr(lib, fn)(d, s, i, fl, ns, nd, cv.scratch(0))
            elif kind == "dpe_pu":
                d = out(cv, nd, ty)
                getattr(lib, fn)(d, s, i, v, ns, nd, cv.scratch(getattr(lib, fn + "_scratch")(ns, nd)))
            elif kind == "dfp_pu":
                d = out(cv, nd, ty)
                getattr(lib, fn)(d, s, i, fl, v, ns, nd, cv.scratch(getattr(lib, fn + "_scratch")(ns, nd)))
        else:
            if kind == "smp_pe":
                lens = f["lens"]; m = lens.size; n = int(lens.sum())
                sd = cv.segd(lens)
                d = out(cv, n, ty)
                getattr(lib, fn)(d, s, i, sd, n, m, cv.scratch(0))
            else:
                ls, ld = f["lens_s"], f["lens_d"]
                m = ls.size
                n_s, n_d = int(ls.sum()), int(ld.sum())
                sds, sdd = cv.segd(ls), cv.segd(ld)
                scr = cv.scratch(getattr(lib, fn + "_scratch")(n_s, m, n_d, m))
                d = out(cv, n_d, ty)
                if kind == "bck_pe":
                    getattr(lib, fn)(d, s, i, sds, n_s, m, sdd, n_d, m, scr)
                elif kind == "fpm_pe":
                    getattr(lib, fn)(d, s, i, fl, sds, n_s, m, sdd, n_d, m, scr)
                elif kind == "bfp_pe":
                    getattr(lib, fn)(d, s, i, fl, sds, n_s, m, sdd, n_d, m, scr)
                elif kind == "dpe_pe":
                    getattr(lib, fn)(d, s, i, v, sds, n_s, m, sdd, n_d, m, scr)
                elif kind == "dfp_pe":
                    getattr(lib, fn)(d, s, i, fl, v, sds, n_s, m, sdd, n_d, m, scr)
        eq(host(d)[:nd], want, fn + case)
    assert lib.cvlhip_check_device_errors() == 0


def test_rank(cv):
    lib = cv.lib
    cases = golden_cases(load_golden("rank"), "")
    for (fn, case), f in sorted(cases.items()):
        s = f["s"]
        n = s.size
        note(fn, case, n)
        d = out(cv, n, "z")
        if fn[4:6] == "lu":
            getattr(lib, fn)(d, dev(cv, s), n, cv.scratch(getattr(lib, fn + "_scratch")(n)))
        else:
            lens = f["lens"]; m = lens.size
            sd = cv.segd(lens)
            getattr(lib, fn)(d, dev(cv, s), sd, n, m, cv.scratch(getattr(lib, fn + "_scratch")(n, m)))
        eq(host(d)[:n], f["d"], fn + case)


def test_misc(cv):
    lib = cv.lib
    cases = golden_cases(load_golden("misc"), "")
    for (fn, case), f in sorted(cases.items()):
        note(fn, case)
        want = f["d"]
        if fn == "ind_luz":
            init, stride = (int(x) for x in f["args"])
            d = out(cv, want.size, "z")
            lib.ind_luz(d, init, stride, want.size, cv.scratch(0))
            eq(host(d)[:want.size], want, fn + case)
            continue
        lens = f["lens"]; m = lens.size; n = int(lens.sum())
        sd = cv.segd(lens)
        ty = fn[6] if len(fn) > 6 else "z"
        if fn == "ind_lez":
            d = out(cv, n, "z")
            lib.ind_lez(d, dev(cv, f["init"]), dev(cv, f["stride"]), sd, n, m, cv.scratch(0))
        elif fn.startswith("dis_ve"):
            d = out(cv, n, ty)
            getattr(lib, fn)(d, dev(cv, f["v"]), sd, n, m, cv.scratch(0))
        elif fn.startswith("ext_ve"):
            d = out(cv, m, ty)
            getattr(lib, fn)(d, dev(cv, f["s"]), dev(cv, f["i"]), sd, n, m, cv.scratch(0))
        elif fn.startswith("rep_ve"):
            d = dev(cv, f["s"])
            getattr(lib, fn)(d, dev(cv, f["i"]), dev(cv, f["v"]), sd, n, m, cv.scratch(0))
        eq(host(d)[:want.size], want, fn + case)


EW1 = {"not_wub", "not_wuz", "int_wud", "int_wub", "dbl_wuz", "boo_wuz", "flr_wud", "cei_wud",
       "trn_wud", "rou_wud", "exp_wud", "log_wud", "sqt_wud", "sin_wud", "cos_wud", "tan_wud",
       "asn_wud", "acs_wud", "atn_wud", "snh_wud", "csh_wud", "tnh_wud"}
TRANSCENDENTAL = {"exp_wud", "log_wud", "sin_wud", "cos_wud", "tan_wud", "asn_wud", "acs_wud",
                  "atn_wud", "snh_wud", "csh_wud", "tnh_wud"}


def test_elementwise(cv):
    lib = cv.lib
    g = load_golden("elwise")
    names = sorted({k.split("/")[0] for k in g})
    for name in names:
        f = {k.split("/")[1]: v for k, v in g.items() if k.split("/")[0] == name}
        want = f["d"]
        n = want.size
        note(name, n)
        d = out(cv, n, "d" if want.dtype == np.float64 else "z")
        if name in EW1:
            getattr(lib, name)(d, dev(cv, f["s"]), n, cv.scratch(0))
        elif name.startswith("sel_wu"):
            getattr(lib, name)(d, dev(cv, f["f"]), dev(cv, f["a"]), dev(cv, f["b"]), n, cv.scratch(0))
        else:
            getattr(lib, name)(d, dev(cv, f["a"]), dev(cv, f["b"]), n, cv.scratch(0))
        got = host(d)[:n]
        if name in TRANSCENDENTAL:
            ulp = np.spacing(np.abs(want))
            assert np.all(np.abs(got - want) <= 4 * ulp), (name, float(np.max(np.abs(got - want) / ulp)))
        else:
            eq(got, want, name)


def test_rand(cv):
    lib = cv.lib
    g = load_golden("rand")
    lib.cvlhip_rand_set_state(0)
    for case in ("first", "second"):
        s = g[f"rnd_wuz/{case}/s"]
        note("rnd_wuz", case)
        d = out(cv, s.size, "z")
        lib.rnd_wuz(d, dev(cv, s), s.size, cv.scratch(0))
        eq(host(d)[:s.size], g[f"rnd_wuz/{case}/d"], "rnd_wuz " + case)
