"""Pins the CPU oracle (oracle/cvl_oracle.c) before it is trusted as the
parity checker: golden vectors from the compiled reference serial CVL, the
known-answer tests of nesl/neslsrc/test.nesl, closed forms / brute force on
tiny inputs, and glibc's own lrand48."""
import ctypes
import subprocess
import sys
from fractions import Fraction

import numpy as np
import pytest

from conftest import golden_cases, load_golden
from oracle import oracle as O
from oracle import refcvl


def eq(a, b):
    a = np.asarray(a); b = np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    if a.dtype == np.float64:
        assert np.array_equal(a.view(np.int64), b.view(np.int64))
    else:
        assert np.array_equal(a, b)


# ---------------------------------------------------------------- golden
def test_golden_scans_reduces():
    g = load_golden("scans")
    cases = golden_cases(g, "")
    assert len(cases) > 300
    for (fn, case), f in cases.items():
        op, kind = fn[:3], fn[4:6]
        ty = fn[6]
        if kind == "su":
            eq(O.scan(op, ty, f["s"]), f["d"])
        elif kind == "ru":
            r = O.reduce(op, ty, f["s"])
            eq(np.array([r], f["r"].dtype), f["r"])
        elif kind == "se":
            eq(O.segscan(op, ty, f["s"], f["lens"]), f["d"])
        elif kind == "re":
            eq(O.segreduce(op, ty, f["s"], f["lens"]), f["d"])


def test_golden_pack():
    cases = golden_cases(load_golden("pack"), "")
    for (fn, case), f in cases.items():
        if fn == "pk1_luv":
            assert O.pk1_luv(f["f"]) == int(f["r"][0])
        elif fn == "pk1_lev":
            eq(O.pk1_lev(f["f"], f["lens"]), f["d"])
        else:
            eq(O.pk2(fn[6], f["s"], f["f"]), f["d"])


def test_golden_permute():
    cases = golden_cases(load_golden("permute"), "")
    assert len(cases) >= 200
    for (fn, case), f in cases.items():
        ty = fn[6]
        kind = fn[:6]
        if kind == "smp_pu":
            got = O.smp_pu(ty, f["s"], f["i"])
        elif kind == "smp_pe":
            got = O.smp_pe(ty, f["s"], f["i"], f["lens"])
        elif kind == "bck_pu":
            got = O.bck_pu(ty, f["s"], f["i"])
        elif kind == "bck_pe":
            got = O.bck_pe(ty, f["s"], f["i"], f["lens_s"], f["lens_d"])
        elif kind == "fpm_pu":
            got = O.fpm_pu(ty, f["s"], f["i"], f["f"], f["d"].size)
        elif kind == "fpm_pe":
            got = O.fpm_pe(ty, f["s"], f["i"], f["f"], f["lens_s"], f["lens_d"])
        elif kind == "bfp_pu":
            got = O.bfp_pu(ty, f["s"], f["i"], f["f"])
        elif kind == "bfp_pe":
            got = O.bfp_pe(ty, f["s"], f["i"], f["f"], f["lens_s"], f["lens_d"])
        elif kind == "dpe_pu":
            got = O.dpe_pu(ty, f["s"], f["i"], f["v"])
        elif kind == "dpe_pe":
            got = O.dpe_pe(ty, f["s"], f["i"], f["v"], f["lens_s"], f["lens_d"])
        elif kind == "dfp_pu":
            got = O.dfp_pu(ty, f["s"], f["i"], f["f"], f["v"])
        elif kind == "dfp_pe":
            got = O.dfp_pe(ty, f["s"], f["i"], f["f"], f["v"], f["lens_s"], f["lens_d"])
        else:
            raise AssertionError(fn)
        eq(got, f["d"])


def test_golden_rank():
    cases = golden_cases(load_golden("rank"), "")
    for (fn, case), f in cases.items():
        up = fn[2] == "u"
        lens = f.get("lens")
        eq(O.rank(f["s"], up, lens), f["d"])


def test_golden_misc():
    cases = golden_cases(load_golden("misc"), "")
    for (fn, case), f in cases.items():
        if fn == "ind_luz":
            init, stride = (int(x) for x in f["args"])
            eq(O.ind_luz(init, stride, f["d"].size), f["d"])
        elif fn == "ind_lez":
            eq(O.ind_lez(f["init"], f["stride"], f["lens"]), f["d"])
        elif fn.startswith("dis_ve"):
            eq(O.dis_ve(fn[6], f["v"], f["lens"]), f["d"])
        elif fn.startswith("ext_ve"):
            eq(O.ext_ve(fn[6], f["s"], f["i"], f["lens"]), f["d"])
        elif fn.startswith("rep_ve"):
            eq(O.rep_ve(fn[6], f["s"], f["i"], f["v"], f["lens"]), f["d"])


EW1_MAP = {"not_wub": "notb", "not_wuz": "notz", "int_wud": "int_d", "int_wub": "int_b",
           "dbl_wuz": "dbl_z", "boo_wuz": "boo_z", "flr_wud": "flr", "cei_wud": "cei",
           "trn_wud": "trn", "rou_wud": "rou", "exp_wud": "exp", "log_wud": "log",
           "sqt_wud": "sqt", "sin_wud": "sin", "cos_wud": "cos", "tan_wud": "tan",
           "asn_wud": "asn", "acs_wud": "acs", "atn_wud": "atn", "snh_wud": "snh",
           "csh_wud": "csh", "tnh_wud": "tnh"}


def test_golden_elwise():
    g = load_golden("elwise")
    names = {k.split("/")[0] for k in g}
    assert len(names) >= 55
    for name in sorted(names):
        f = {k.split("/")[1]: v for k, v in g.items() if k.split("/")[0] == name}
        if name in EW1_MAP:
            eq(O.ew1(EW1_MAP[name], f["s"]), f["d"])
        elif name.startswith("sel_wu"):
            eq(O.sel(name[6], f["f"], f["a"], f["b"]), f["d"])
        else:
            eq(O.ew2(name[:3], name[6], f["a"], f["b"]), f["d"])


def test_golden_rand_and_lrand48():
    g = load_golden("rand")
    rng = O.Lrand48(0)       # glibc starts an unseeded process at state 0
    eq(rng.rnd_wuz(g["rnd_wuz/first/s"]), g["rnd_wuz/first/d"])
    eq(rng.rnd_wuz(g["rnd_wuz/second/s"]), g["rnd_wuz/second/d"])


def test_lrand48_matches_glibc_fresh_process():
    """glibc's lrand48 in a fresh process, unseeded and after srand48."""
    code = ("import ctypes; c=ctypes.CDLL('libc.so.6'); c.lrand48.restype=ctypes.c_long;"
            "a=[c.lrand48() for _ in range(50)]; c.srand48(ctypes.c_long(12345));"
            "b=[c.lrand48() for _ in range(50)]; print(a); print(b)")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         check=True).stdout.splitlines()
    a = eval(out[0]); b = eval(out[1])  # noqa: S307 (our own subprocess output)
    r = O.Lrand48(0)
    assert [r.next() for _ in range(50)] == a
    r = O.Lrand48.seeded(12345)
    assert [r.next() for _ in range(50)] == b


def test_reference_rank_word_order():
    """The as-is x86-64 build of the reference takes rank.c's big-endian
    branch and fails test.nesl's float rank KAT; the oracle follows the
    little-endian branch, which passes it."""
    g = load_golden("rand")
    s = np.array([10.0, -5.0, 0.0, 10.0, -10.0, 100.0, 10.0])
    if "asis_rku_lud/kat/d" in g:
        assert not np.array_equal(g["asis_rku_lud/kat/d"], [3, 1, 2, 4, 0, 6, 5])
    eq(O.rank(s, True), np.array([3, 1, 2, 4, 0, 6, 5], np.int32))


# ------------------------------------------------------ test.nesl KATs
def test_nesl_kats():
    # nesl/neslsrc/test.nesl:62-120
    eq(O.scan("add", "z", [2, -4, 3, 3]), np.int32([0, 2, -2, 1]))
    eq(O.scan("add", "d", [2.0, -3.0, 3.0]), np.float64([0.0, 2.0, -1.0]))
    eq(O.scan("mul", "z", [2, -3, 3]), np.int32([1, 2, -6]))
    eq(O.scan("mul", "d", [2.0, -3.0, 3.0]), np.float64([1.0, 2.0, -6.0]))
    assert O.reduce("max", "z", [-5, 23, -100, 7, 0, -2]) == 23
    assert O.reduce("min", "z", [-5, 23, -100, 7000, 0, -2]) == -100
    v = np.int32([-5, 8, 1])
    eq(np.maximum(v, O.scan("max", "z", v)), np.int32([-5, 8, 8]))
    v = np.int32([5, -8, -1])
    eq(np.minimum(v, O.scan("min", "z", v)), np.int32([5, -8, -8]))
    v = np.float64([-5.0, 8.0, 1.0])
    eq(np.maximum(v, O.scan("max", "d", v)), np.float64([-5.0, 8.0, 8.0]))
    T, F = 1, 0
    eq(O.scan("ior", "b", [F, F, F, T, F, F, F]), np.int32([F, F, F, F, T, T, T]))
    eq(O.scan("and", "b", [T, T, T, F, T, F, T]), np.int32([T, T, T, T, F, F, F]))
    eq(O.scan("xor", "b", [F, T, F, T, F, F, F]), np.int32([F, F, T, T, F, F, F]))
    eq(O.scan("ior", "z", [6, 1, 32768, 63, 32704, 99]), np.int32([0, 6, 7, 32775, 32831, 65535]))
    eq(O.scan("and", "z", [32767, 29, 11, 9, 0, 65535]), np.int32([-1, 32767, 29, 9, 9, 0]))
    eq(O.scan("xor", "z", [0, 0, 0, 32767, 32767, 32767, 3, 0, 3]),
       np.int32([0, 0, 0, 0, 32767, 0, 32767, 32764, 32764]))
    # permute / get / put / pack (test.nesl:122-173)
    s = np.frombuffer(b"mowtab", np.uint8).astype(np.int32)
    eq(O.smp_pu("z", s, [2, 1, 0, 5, 4, 3]), np.frombuffer(b"wombat", np.uint8).astype(np.int32))
    src = np.frombuffer(b"testing it", np.uint8).astype(np.int32)
    eq(O.bck_pu("z", src, [3, 4, 1]), np.frombuffer(b"tie", np.uint8).astype(np.int32))
    eq(O.dpe_pu("z", s, [2, 1, 0, 6, 5, 4], src),
       np.frombuffer(b"womtbat it", np.uint8).astype(np.int32))
    h = np.frombuffer(b"the horse", np.uint8).astype(np.int32)
    eq(O.pk2("z", h, [T, F, T, T, F, F, T, T, F]), np.frombuffer(b"te rs", np.uint8).astype(np.int32))
    eq(O.pk2("b", [T, F, T, F, T, T, T, F, F], [T, F, T, T, F, F, T, T, F]), np.int32([T, T, F, T, F]))
    eq(O.pk2("d", np.arange(8.0), [T, F, T, F, T, T, F, T]), np.float64([0, 2, 4, 5, 7]))
    eq(O.smp_pu("b", [F, F, T, T, T, T], [2, 4, 0, 1, 3, 5]), np.int32([T, T, F, T, F, T]))
    # rank / rank_down (test.nesl:627-636)
    k = [10, -5, 0, 10, -10, 100, 10]
    eq(O.rank(np.int32(k), True), np.int32([3, 1, 2, 4, 0, 6, 5]))
    eq(O.rank(np.int32(k), False), np.int32([1, 5, 4, 2, 6, 0, 3]))
    eq(O.rank(np.float64(k), True), np.int32([3, 1, 2, 4, 0, 6, 5]))
    eq(O.rank(np.float64(k), False), np.int32([1, 5, 4, 2, 6, 0, 3]))


def test_nesl_bigtest_identities():
    """bigtest (test.nesl:436-481) at len 50000 and its segmented form."""
    n = 50000
    eq(O.scan("add", "z", np.ones(n, np.int32)), O.ind_luz(0, 1, n))
    idx = O.ind_luz(0, 1, n)
    rev = (n - 1 - idx).astype(np.int32)
    eq(O.smp_pu("z", idx, rev), rev)
    assert O.reduce("add", "z", np.full(n, 5, np.int32)) == 5 * n
    flags = (idx < n // 2).astype(np.int32)
    eq(O.pk2("z", idx, flags), O.ind_luz(0, 1, n // 2))
    lens = np.int32([n // 6, n // 2, n // 3])
    ones = np.ones(int(lens.sum()), np.int32)
    eq(O.segscan("add", "z", ones, lens),
       O.ind_lez(np.zeros(3, np.int32), np.ones(3, np.int32), lens))


# ------------------------------------------------ closed forms, brute force
@pytest.mark.parametrize("seed", range(5))
def test_closed_forms_small(seed):
    rng = np.random.default_rng(seed)
    n = 37
    s = rng.integers(-2 ** 31, 2 ** 31, n, dtype=np.int64).astype(np.int32)
    pre = np.concatenate([[0], np.cumsum(s.astype(np.int64))[:-1]])
    eq(O.scan("add", "z", s), ((pre + 2 ** 31) % 2 ** 32 - 2 ** 31).astype(np.int32))
    mx = np.concatenate([[np.iinfo(np.int32).min], np.maximum.accumulate(s)[:-1]]).astype(np.int32)
    eq(O.scan("max", "z", s), mx)
    mn = np.concatenate([[np.iinfo(np.int32).max], np.minimum.accumulate(s)[:-1]]).astype(np.int32)
    eq(O.scan("min", "z", s), mn)
    xr = np.concatenate([[0], np.bitwise_xor.accumulate(s)[:-1]]).astype(np.int32)
    eq(O.scan("xor", "z", s), xr)
    # wrapped product
    p, out = 1, []
    for x in s:
        out.append(((p + 2 ** 31) % 2 ** 32) - 2 ** 31)
        p = (p * int(x)) % 2 ** 32
    eq(O.scan("mul", "z", s), np.array(out, np.int64).astype(np.int32))
    # identities on doubles are the float-rounded 1e36
    assert O.scan("max", "d", [3.0])[0] == float(np.float32(-1e36))
    assert O.scan("min", "d", [3.0])[0] == float(np.float32(1e36))
    # segmented scan == per-segment scan
    lens = np.int32([5, 0, 12, 1, 19])
    seg = O.segscan("add", "z", s, lens)
    o = np.concatenate([[0], np.cumsum(lens)])
    for g in range(lens.size):
        eq(seg[o[g]:o[g + 1]], O.scan("add", "z", s[o[g]:o[g + 1]]))
    r = O.segreduce("max", "z", s, lens)
    assert r[1] == np.iinfo(np.int32).min


@pytest.mark.parametrize("seed", range(4))
def test_rank_is_stable_sort(seed):
    rng = np.random.default_rng(seed)
    s = rng.integers(-5, 5, 200).astype(np.int32)
    r = O.rank(s, True)
    order = np.argsort(s, kind="stable")
    exp = np.empty_like(r); exp[order] = np.arange(s.size)
    eq(r, exp)
    r = O.rank(s, False)
    order = np.argsort(-s.astype(np.int64), kind="stable")
    exp[order] = np.arange(s.size)
    eq(r, exp)
    lens = np.int32([50, 0, 100, 50])
    r = O.rank(s, True, lens)
    o = np.concatenate([[0], np.cumsum(lens)])
    for g in range(4):
        eq(r[o[g]:o[g + 1]], O.rank(s[o[g]:o[g + 1]], True))
    d = rng.uniform(-1, 1, 200)
    order = np.argsort(d, kind="stable")
    exp[order] = np.arange(d.size)
    eq(O.rank(d, True), exp)


def test_pack_permute_invariants():
    rng = np.random.default_rng(7)
    n = 5000
    s = rng.integers(0, 100, n).astype(np.int32)
    f = (rng.random(n) < 0.3).astype(np.int32)
    assert O.pk2("z", s, f).size == O.reduce("add", "z", f)
    p = rng.permutation(n).astype(np.int32)
    inv = np.empty_like(p); inv[p] = np.arange(n)
    eq(O.smp_pu("z", O.smp_pu("z", s, p), inv), s)
    eq(O.bck_pu("z", O.smp_pu("z", s, p), p), s)


def test_neumaier_exact_on_small():
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, 300) * 10.0 ** rng.integers(-8, 8, 300)
    lens = np.int32([100, 0, 200])
    out, absp = O.segscan_add_neumaier(x, lens)
    o = np.concatenate([[0], np.cumsum(lens)])
    for g in range(3):
        acc = Fraction(0)
        for k in range(o[g], o[g + 1]):
            exact = float(acc)
            assert abs(out[k] - exact) <= 2 * 2.0 ** -53 * max(absp[k], 1e-300)
            acc += Fraction(x[k])
    r, _ = O.segreduce_add_neumaier(x, lens)
    assert r[1] == 0.0


def test_splitmix_reference_values():
    # splitmix64 of seed 0, first three outputs (published test values of
    # the generator: 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f)
    u = O.splitmix_u32(3, 0)
    assert [int(v) for v in u] == [0xe220a839, 0x6e789e6a, 0x06c45d18]


# ------------------------------------- oracle vs compiled reference (optional)
@pytest.mark.skipif(not refcvl.available(), reason="reference serial CVL not built here")
def test_oracle_vs_reference_random():
    rng = np.random.default_rng(11)
    n = 20000
    s = rng.integers(-2 ** 31, 2 ** 31, n, dtype=np.int64).astype(np.int32)
    lens = rng.geometric(1 / 50, 2000).astype(np.int32) - 1
    lens[-1] += n - lens.sum() if lens.sum() < n else 0
    while lens.sum() > n:
        lens = lens[:-1]
    lens = np.append(lens, n - lens.sum()).astype(np.int32)
    m = lens.size
    for op in ("add", "mul", "min", "max", "and", "ior", "xor"):
        d = np.zeros(n, np.int32)
        refcvl.call(f"{op}_sez", d, s, lens, n, m, refcvl.scratch_for(1))
        eq(O.segscan(op, "z", s, lens), d)
    r = np.zeros(n, np.int32)
    refcvl.call("rku_lez", r, s.copy(), lens, n, m, refcvl.scratch_for(n + 1))
    eq(O.rank(s, True, lens), r)
    x = rng.uniform(-1, 1, n)
    d = np.zeros(n)
    refcvl.call("add_sed", d, x, lens, n, m, refcvl.scratch_for(1))
    eq(O.segscan("add", "d", x, lens), d)
