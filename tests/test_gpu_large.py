"""Parity at the benchmark sizes: 2^28 int32 for scans, reduce, pack and
permute, 2^26 keys for rank, 2^26 doubles in Zipf(1.5) segments for the
irregular float case.  Integer results are compared bit-exactly with the C
oracle or with exact numpy closed forms (stable argsort, boolean
compaction, inverse permutation); double +-scans against the compensated
oracle with the tolerance 2(ceil(log2 L)+2) 2^-53 sum|x|."""
import math
import time

import numpy as np
import pytest

from gpu_util import eq, host
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = 1 << 28
NR = 1 << 26
SEED = 0x4E45534C00000002


@pytest.fixture(scope="module")
def cv():
    from neslgpu_amd.cvl import Cvl
    return Cvl()


def note(*a):
    print(time.strftime("%H:%M:%S"), *a, flush=True)


def geometric(n, mean, seed):
    rng = np.random.default_rng(seed)
    lens = rng.geometric(1.0 / mean, size=n // mean + 4096)
    cs = np.cumsum(lens)
    k = int(np.searchsorted(cs, n))
    lens = lens[:k + 1]
    lens[-1] -= int(cs[k] - n)
    return lens[lens > 0].astype(np.int32)


def zipf_lengths(n, alpha, lmax, seed):
    rng = np.random.default_rng(seed)
    out, tot = [], 0
    while tot < n:
        z = rng.zipf(alpha, size=1 << 20)
        z = np.minimum(z, lmax)
        out.append(z)
        tot += int(z.sum())
    lens = np.concatenate(out)
    cs = np.cumsum(lens)
    k = int(np.searchsorted(cs, n))
    lens = lens[:k + 1].copy()
    lens[-1] -= int(cs[k] - n)
    return lens[lens > 0].astype(np.int32)


@pytest.fixture(scope="module")
def data(cv):
    import torch
    L = cv.lib
    s = torch.empty(N, dtype=torch.int32, device=cv.device)
    L.cvlhip_fill_splitmix_u32(s, N, SEED)
    s_h = O.splitmix_u32(N, SEED).view(np.int32)
    eq(host(s[:1 << 20]), s_h[:1 << 20], "device and host splitmix streams")
    return s, s_h


def test_add_suz_full(cv, data):
    import torch
    s, s_h = data
    note("add_suz 2^28")
    d = torch.empty_like(s)
    cv.lib.add_suz(d, s, N, cv.scratch(cv.lib.add_suz_scratch(N)))
    eq(host(d), O.scan("add", "z", s_h), "add_suz 2^28")
    note("add_suz in place")
    d.copy_(s)
    cv.lib.add_suz(d, d, N, cv.scratch(cv.lib.add_suz_scratch(N)))
    eq(host(d[-(1 << 20):]), O.scan("add", "z", s_h)[-(1 << 20):], "add_suz in place (tail)")


def test_max_ruz_full(cv, data):
    s, s_h = data
    note("max_ruz 2^28")
    r = cv.lib.max_ruz(s, N, cv.scratch(cv.lib.max_ruz_scratch(N)))
    assert r == int(s_h.max())
    r = cv.lib.min_ruz(s, N, cv.scratch(cv.lib.min_ruz_scratch(N)))
    assert r == int(s_h.min())
    r = cv.lib.add_ruz(s, N, cv.scratch(cv.lib.add_ruz_scratch(N)))
    assert r == int(((int(s_h.astype(np.int64).sum()) + 2 ** 31) % 2 ** 32) - 2 ** 31)


def test_add_sez_full(cv, data):
    import torch
    s, s_h = data
    lens = geometric(N, 256, 7)
    note("add_sez 2^28, m =", lens.size)
    sd = cv.segd(lens)
    d = torch.empty_like(s)
    cv.lib.add_sez(d, s, sd, N, lens.size, cv.scratch(cv.lib.add_sez_scratch(N, lens.size)))
    eq(host(d), O.segscan("add", "z", s_h, lens), "add_sez 2^28")
    note("max_rez 2^28")
    r = torch.empty(lens.size, dtype=torch.int32, device=cv.device)
    cv.lib.max_rez(r, s, sd, N, lens.size, cv.scratch(cv.lib.max_rez_scratch(N, lens.size)))
    eq(host(r), O.segreduce("max", "z", s_h, lens), "max_rez 2^28")


def test_pack_full(cv, data):
    import torch
    s, s_h = data
    f = (s < 0).to(torch.int32)
    f_h = (s_h < 0)
    kept = int(f_h.sum())
    note("pk2_luz 2^28, kept", kept)
    assert cv.lib.pk1_luv(f, N, cv.scratch(cv.lib.pk1_luv_scratch(N))) == kept
    d = torch.empty(kept, dtype=torch.int32, device=cv.device)
    cv.lib.pk2_luz(d, s, f, N, kept, cv.scratch(cv.lib.pk2_luz_scratch(N)))
    eq(host(d), s_h[f_h], "pk2_luz 2^28")
    assert cv.lib.cvlhip_check_device_errors() == 0


def test_permute_full(cv, data):
    import torch
    s, s_h = data
    g = torch.Generator(device=cv.device)
    g.manual_seed(5)
    p = torch.randperm(N, device=cv.device, generator=g, dtype=torch.int64).to(torch.int32)
    note("smp_puz 2^28")
    d = torch.empty_like(s)
    cv.lib.smp_puz(d, s, p, N, cv.scratch(cv.lib.smp_puz_scratch(N)))
    p_h = host(p)
    want = np.empty_like(s_h)
    want[p_h] = s_h
    eq(host(d), want, "smp_puz 2^28")
    note("bck_puz 2^28 (inverse)")
    b = torch.empty_like(s)
    cv.lib.bck_puz(b, d, p, N, N, cv.scratch(cv.lib.bck_puz_scratch(N, N)))
    eq(host(b), s_h, "bck_puz o smp_puz == id")
    assert cv.lib.cvlhip_check_device_errors() == 0


def test_rank_full(cv):
    import torch
    k = torch.empty(NR, dtype=torch.int32, device=cv.device)
    cv.lib.cvlhip_fill_splitmix_u32(k, NR, SEED + 1)
    k_h = O.splitmix_u32(NR, SEED + 1).view(np.int32)
    note("rku_luz 2^26")
    r = torch.empty(NR, dtype=torch.int32, device=cv.device)
    cv.lib.rku_luz(r, k, NR, cv.scratch(cv.lib.rku_luz_scratch(NR)))
    order = np.argsort(k_h, kind="stable")
    want = np.empty(NR, np.int32)
    want[order] = np.arange(NR, dtype=np.int32)
    eq(host(r), want, "rku_luz 2^26")
    note("rkd_luz 2^26")
    cv.lib.rkd_luz(r, k, NR, cv.scratch(cv.lib.rkd_luz_scratch(NR)))
    order = np.argsort(-k_h.astype(np.int64), kind="stable")
    want[order] = np.arange(NR, dtype=np.int32)
    eq(host(r), want, "rkd_luz 2^26")
    assert cv.lib.cvlhip_check_device_errors() == 0


def test_irregular_f64_full(cv):
    """2^26 doubles in [-1,1], Zipf(1.5) lengths in [1, 2^20]: add_sed within
    the compensated-oracle tolerance, max_red bit-exact."""
    import torch
    n = 1 << 26
    lens = zipf_lengths(n, 1.5, 1 << 20, 11)
    m = lens.size
    note("irregular f64: n", n, "m", m, "max len", int(lens.max()))
    x = torch.empty(n, dtype=torch.float64, device=cv.device)
    cv.lib.cvlhip_fill_splitmix_f64(x, n, SEED + 2, -1.0, 1.0)
    x_h = O.splitmix_f64(n, SEED + 2, -1.0, 1.0)
    eq(host(x[:1 << 16]), x_h[:1 << 16], "device and host f64 streams")
    sd = cv.segd(lens)
    d = torch.empty_like(x)
    cv.lib.add_sed(d, x, sd, n, m, cv.scratch(cv.lib.add_sed_scratch(n, m)))
    ref, absp = O.segscan_add_neumaier(x_h, lens)
    seglen = np.repeat(lens.astype(np.int64), lens)
    tol = 2.0 * (np.ceil(np.log2(np.maximum(seglen, 2))) + 2) * 2.0 ** -53 * absp
    err = np.abs(host(d) - ref)
    assert np.all(err <= tol), float((err - tol).max())
    note("max_red")
    r = torch.empty(m, dtype=torch.float64, device=cv.device)
    cv.lib.max_red(r, x, sd, n, m, cv.scratch(cv.lib.max_red_scratch(n, m)))
    eq(host(r), O.segreduce("max", "d", x_h, lens), "max_red")
