"""Segment descriptors (mke_fov) read back section by section and compared
with their definition, for both construction paths: one workgroup for small
descriptors (m and the bitmap's word count at most 4096), the two look-back
kernels otherwise.

Layout (byte offsets, each section 256-byte aligned): header (n, m as
int64) at 0; lens[m] at 256; offs[m+1] (exclusive sums, offs[m] = n); nz
(ids of the non-empty segments, in order); wcnt[W+1] (segment heads in the
32-element words before each word); heads[W+8] (bit k of word k/32 set when
a non-empty segment starts at element k).  W = ceil(n / 32).  The lengths
are the serial CVL's whole descriptor (nesl/cvl/serial/facilt.c:118-144)."""
import numpy as np
import pytest

from gpu_util import host

pytestmark = pytest.mark.gpu


def up256(b):
    return (b + 255) & ~255


def layout(n, m):
    words = (n + 31) // 32
    lens = 256
    offs = up256(lens + 4 * m)
    nz = up256(offs + 4 * (m + 1))
    wcnt = up256(nz + 4 * m)
    heads = up256(wcnt + 4 * (words + 1))
    return words, lens, offs, nz, wcnt, heads


def want(lens):
    lens = np.asarray(lens, np.int64)
    n, m = int(lens.sum()), lens.size
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    nz = np.nonzero(lens > 0)[0]
    words = (n + 31) // 32
    heads = np.zeros(words + 8, np.uint32)
    starts = offs[:-1][lens > 0]
    np.bitwise_or.at(heads, starts // 32, (np.uint32(1) << (starts % 32).astype(np.uint32)))
    pc = np.array([bin(int(h)).count("1") for h in heads[:words]], np.int64)
    wcnt = np.concatenate([[0], np.cumsum(pc)])
    return n, m, offs, nz, words, heads, wcnt


def shapes():
    rng = np.random.default_rng(21)

    def split(n, m, empty=0.0):
        cuts = np.sort(rng.integers(0, n + 1, m - 1))
        lens = np.diff(np.concatenate([[0], cuts, [n]]))
        if empty:
            lens[rng.random(m) < empty] = 0
        return lens.astype(np.int32)
    return [
        ("small", split(5000, 100, 0.1)),
        ("small_edge", split(4096 * 32, 4096, 0.1)),     # the largest one-workgroup descriptor
        ("few_long", split(500_000, 3000, 0.05)),        # look-back kernels: bitmap above 4096 words
        ("mid_tiles", split(700_000, 20_000, 0.2)),      # several tiles of lengths and of words, empties
        ("mid_edge", split(1 << 20, 1 << 15, 0.1)),      # 8 tiles of each, exactly
        ("many_short", split(120_000, 50_000, 0.2)),     # more than 4096 lengths
        ("large_long", split(1 << 21, 70_000)),
        ("large_many", split(300_000, 200_000, 0.3)),
        ("all_empty", np.zeros(5000, np.int32)),
        ("singletons", np.ones(100_003, np.int32)),     # every bit of every bitmap word set
        ("pairs_empty", np.tile(np.array([1, 0, 2], np.int32), 40_000)),
        ("empty_ends", np.concatenate([[0, 0], split(200_000, 9000), [0]]).astype(np.int32)),
    ]


@pytest.mark.parametrize("name,lens", shapes(), ids=[s[0] for s in shapes()])
def test_descriptor_sections(cv, name, lens):
    n, m, offs, nz, words, heads, wcnt = want(lens)
    print("segd", name, n, m, flush=True)
    sd = cv.segd(lens)
    raw = host(sd.buf).view(np.uint8)
    W, o_lens, o_offs, o_nz, o_wcnt, o_heads = layout(n, m)
    assert W == words
    hdr = raw[:16].view(np.int64)
    assert hdr[0] == n and hdr[1] == m
    i32 = lambda off, k: raw[off:off + 4 * k].view(np.int32)
    assert np.array_equal(i32(o_lens, m), lens)
    assert np.array_equal(i32(o_offs, m + 1), offs)
    assert np.array_equal(i32(o_nz, nz.size), nz)
    assert np.array_equal(i32(o_wcnt, words + 1), wcnt)
    assert np.array_equal(raw[o_heads:o_heads + 4 * (words + 8)].view(np.uint32), heads)
    assert cv.lib.cvlhip_check_device_errors() == 0


@pytest.fixture(scope="module")
def cv():
    from neslgpu_amd.cvl import Cvl
    return Cvl()
