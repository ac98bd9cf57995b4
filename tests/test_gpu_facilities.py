"""CVL facilities and scalar entry points on the GPU (ref
nesl/cvl/serial/facilt.c:92-318, vprims.c:207-344, elwise.c:291-298):
sizes, the arena, pointer arithmetic in 8-byte units, the overlap-safe
move, host <-> device copies, distribute / extract / replace on whole
vectors, copies, the timer and the RNG seed.

mov_fov must behave like memmove (facilt.c:161-183 walks forward when the
destination is below the source and backward otherwise), whatever the
overlap; it is checked against numpy on the same bytes."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cv():
    from neslgpu_amd.cvl import Cvl
    return Cvl()


def test_sizes_cover_the_data(cv):
    L = cv.lib
    for n in (0, 1, 2, 3, 1000, 12345, 1 << 20):
        assert L.siz_foz(n) * 8 >= 4 * n and L.siz_fob(n) * 8 >= 4 * n
        assert L.siz_fod(n) * 8 >= 8 * n
        for m in (1, 7, n // 3 + 1):
            words = (n + 31) // 32
            assert L.siz_fos(n, m) * 8 >= 256 + 4 * m + 4 * (m + 1) + 4 * m + 4 * (words + 1) + 4 * (words + 8)


def test_pointer_arithmetic_counts_units(cv):
    L = cv.lib
    p = L.alo_foz(1000)
    assert p
    q = L.add_fov(p, 37)
    assert q - p == 37 * 8
    assert L.sub_fov(q, p) == 37 and L.sub_fov(p, q) == 37     # facilt.c:207-216: distance
    assert L.cmp_fov(q, p) == 37 and L.cmp_fov(p, q) == -37    # facilt.c:226-230: signed
    assert L.eql_fov(p, p) and not L.eql_fov(p, q)
    L.fre_fov(p)


@pytest.mark.parametrize("src,dst,units", [(0, 3, 100), (3, 0, 100), (0, 1, 5000), (1, 0, 5000),
                                           (10, 10, 50), (0, 400, 300), (0, 7, 300_000), (7, 0, 300_000),
                                           (100, 99, 1)])
def test_mov_fov_is_memmove(cv, src, dst, units):
    import torch
    L = cv.lib
    total = max(src, dst) + units + 16
    init = np.arange(2 * total, dtype=np.int32) * 7 + 3          # two ints per 8-byte unit
    arena = torch.from_numpy(init.copy()).to(cv.device)
    p = arena.data_ptr()
    print("mov_fov", src, dst, units, flush=True)
    L.mov_fov(L.add_fov(p, dst), L.add_fov(p, src), units, None)
    want = init.copy()
    want[2 * dst:2 * (dst + units)] = init[2 * src:2 * (src + units)].copy()
    assert np.array_equal(host(arena), want)


@pytest.mark.parametrize("ty", ["z", "b", "d"])
def test_host_copies_and_scalars(cv, ty):
    import torch
    L = cv.lib
    rng = np.random.default_rng(3)
    n = 10_007
    dt = np.float64 if ty == "d" else np.int32
    x = rng.uniform(-5, 5, n) if ty == "d" else (rng.integers(0, 2, n) if ty == "b" else
                                                rng.integers(-2 ** 31, 2 ** 31, n)).astype(np.int32)
    x = x.astype(dt)
    d = torch.empty(n, dtype=torch.float64 if ty == "d" else torch.int32, device=cv.device)
    getattr(L, f"c2v_fu{ty}")(d, x, n, None)                   # host -> device
    back = np.empty_like(x)
    getattr(L, f"v2c_fu{ty}")(back, d, n, None)                # device -> host
    assert np.array_equal(back, x)
    e = torch.empty_like(d)
    getattr(L, f"cpy_wu{ty}")(e, d, n, None)
    assert np.array_equal(host(e), x)
    for i in (0, 1, n // 2, n - 1):                            # ext_vu*: one element to the host
        assert getattr(L, f"ext_vu{ty}")(d, i, n, None) == x[i]
    v = 1 if ty == "b" else (2.5 if ty == "d" else -77)
    getattr(L, f"rep_vu{ty}")(d, 17, v, n, None)               # rep_vu*: one element replaced
    x2 = x.copy()
    x2[17] = v
    assert np.array_equal(host(d), x2)
    getattr(L, f"dis_vu{ty}")(d, v, n, None)                   # dis_vu*: every element
    assert np.array_equal(host(d), np.full(n, v, dtype=dt))


def test_descriptor_copy(cv):
    import torch
    L = cv.lib
    lens = np.array([3, 0, 5, 1, 9], np.int32)
    sd = cv.segd(lens)
    n, m = int(lens.sum()), lens.size
    units = L.siz_fos(n, m)
    dst = torch.zeros(units * 2 + 64, dtype=torch.int32, device=cv.device)
    L.cpy_wus(dst, sd, n, m, None)
    out = np.empty(m, np.int32)
    lo = L.cvlhip_segd_lengths_offset() // 4
    assert np.array_equal(host(dst)[lo:lo + m], lens)
    want = torch.empty(m, dtype=torch.int32, device=cv.device)
    L.len_fos(want, dst, n, m, None)                           # the copy is a whole descriptor
    L.v2c_fuz(out, want, m, None)
    assert np.array_equal(out, lens)


def test_timer_and_seed(cv):
    L = cv.lib
    t0 = (C.c_byte * 64)()
    t1 = (C.c_byte * 64)()
    L.tgt_fos(C.addressof(t0))
    L.tgt_fos(C.addressof(t1))
    assert L.tdf_fos(C.addressof(t1), C.addressof(t0)) >= 0.0
    # rnd_foz does not touch the lrand48 state behind RAND (facilt.c:307-311)
    L.cvlhip_rand_set_state(12345)
    L.rnd_foz(99)
    assert L.cvlhip_rand_get_state() == 12345
    L.cvlhip_rand_set_state(0)
