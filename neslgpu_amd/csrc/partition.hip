// partition.hip -- permute by a permutation without per-element scattered
// stores (the engine behind smp_pu*/smp_pe*, ref nesl/cvl/serial/
// vprims.c:356-405, and the final inversion of the radix rank).
//
// On MI355X a store whose 64 lanes hit 64 different cache lines costs ~64
// line transactions whatever the locality, which caps a naive scatter at
// ~45 G elements/s.  Here every store is coalesced:
//
//   partition passes  Destination space is cut into blocks of kBlk elements.
//                     Because the indices form a permutation, destination
//                     block c receives exactly its size, so pairs (index,
//                     value) can be bucketed by block without a histogram
//                     pass: a tile counts its elements per bucket in LDS,
//                     claims a slot range per bucket with one atomicAdd,
//                     stages the pairs bucket by bucket in LDS and writes
//                     each bucket's run contiguously.  One pass splits a
//                     region into <= 256 children; two passes take 2^28
//                     elements to 8192-element blocks.
//   block scatter     One workgroup per destination block scatters its
//                     pairs into LDS and writes the block out in order.
//
// Pairs whose index is out of range, or that overflow their bucket (the
// indices were not a permutation), are dropped and raise the device error
// word; destination positions nobody wrote keep their old contents.
#include "cvl_hip.hpp"
#include "segd.hpp"
#include "partition.hpp"
#include "colscan.hpp"
#include "../../include/cvl.h"

namespace cvlhip {

constexpr int kBlkLog = 13;
constexpr int64_t kBlk = 1 << kBlkLog;         // final destination block (LDS-resident)
#ifndef CVLHIP_PART_ITEMS
#define CVLHIP_PART_ITEMS 8
#endif
constexpr int kPartThreads = 512;
constexpr int kPartItems = CVLHIP_PART_ITEMS;
constexpr int kPartTile = kPartThreads * kPartItems;   // pairs per tile
constexpr int kMaxFan = 256;                   // children per region per pass

// Where the destination index and the value of input element k come from
// on the first pass: the index is p[k] (plain permute) or offs[seg(k)] + p[k]
// (segmented permute); the value is s[k], or -- for inverting a sort order
// -- k itself, optionally relative to the start of k's segment in `vseg`.
struct IndexSource {
    const int *p;
    SegdView seg;
    bool segmented;
    int vmode;        // 0: s[k]; 1: k; 2: k - offs[seg_of(k)] in vseg
    SegdView vseg;
    __device__ __forceinline__ int64_t at(int64_t k) const {
        int64_t q = p[k];
        if (segmented) q += seg.offs[seg.seg_of(k)];
        return q;
    }
    template <class T>
    __device__ __forceinline__ T value(const T *__restrict__ s, int64_t k) const {
        if (vmode == 0) return s[k];
        if (vmode == 1) return (T)k;
        return (T)(k - vseg.offs[vseg.seg_of(k)]);
    }
};

// Intermediate (index, value) pairs between passes.  Four-byte values are
// interleaved with their index in one 8-byte word, so a bucket's run is one
// contiguous stream of 8-byte stores (a run of 16 pairs fills a 128-byte
// line) instead of two half-as-long streams; wider values keep two arrays.
template <class T, bool PACKED = sizeof(T) == 4>
struct Pairs {
    int *idx;
    T *val;
    __device__ __forceinline__ void load(int64_t e, int &q, T &v) const {
        q = idx[e];
        v = val[e];
    }
    __device__ __forceinline__ void store(int64_t e, int q, T v) const {
        idx[e] = q;
        val[e] = v;
    }
};
template <class T>
struct Pairs<T, true> {
    uint2 *p;
    __device__ __forceinline__ void load(int64_t e, int &q, T &v) const {
        const uint2 w = p[e];
        q = (int)w.x;
        v = __builtin_bit_cast(T, w.y);
    }
    __device__ __forceinline__ void store(int64_t e, int q, T v) const {
        p[e] = make_uint2((unsigned)q, __builtin_bit_cast(unsigned, v));
    }
};

// One partition pass.  Region r = [r*psize, min((r+1)*psize, n)) of the
// input is split into children of csize elements (csize divides psize,
// psize/csize <= 256).  Block b handles tile (b % tiles_per_region) of
// region (b / tiles_per_region).  Child c's pairs land in
// [c*csize, c*csize + cnt[c]) of the output.
template <class T, bool FIRST>
__global__ __launch_bounds__(kPartThreads) void k_partition(IndexSource src, Pairs<T> in,
                                                            const T *__restrict__ in_val, Pairs<T> out,
                                                            int64_t n, int plog, int clog,
                                                            int64_t tiles_per_region,
                                                            const unsigned *__restrict__ prev_cnt,
                                                            unsigned *__restrict__ cnt,
                                                            const unsigned *__restrict__ tile_ofs,
                                                            unsigned *__restrict__ err) {
    __shared__ unsigned s_hist[kMaxFan];
    __shared__ unsigned s_start[kMaxFan];
    __shared__ unsigned s_gofs[kMaxFan];
    __shared__ int s_idx[kPartTile];
    __shared__ T s_val[kPartTile];
    __shared__ unsigned char s_dig[kPartTile];

    const int tid = threadIdx.x;
    const int64_t bt = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t region = bt / tiles_per_region;
    const int64_t chunk = bt % tiles_per_region;
    const int64_t rlo = region << plog;
    int64_t rhi = (rlo + ((int64_t)1 << plog)) < n ? rlo + ((int64_t)1 << plog) : n;
    if (!FIRST && rlo + (int64_t)prev_cnt[region] < rhi) rhi = rlo + prev_cnt[region];   // pairs that landed
    const int64_t lo = rlo + chunk * kPartTile;
    const int64_t first_child = rlo >> clog;

    for (int i = tid; i < kMaxFan; i += kPartThreads) s_hist[i] = 0;
    __syncthreads();

    int q[kPartItems];
    T v[kPartItems];
    unsigned rk[kPartItems];
    int dg[kPartItems];
#pragma unroll
    for (int j = 0; j < kPartItems; j++) {
        const int64_t e = lo + (int64_t)j * kPartThreads + tid;
        dg[j] = -1;
        if (e < rhi) {
            int64_t qq;
            if constexpr (FIRST) {
                qq = src.at(e);
                v[j] = src.value(in_val, e);
            } else {
                int qi;
                in.load(e, qi, v[j]);
                qq = qi;
            }
            const int64_t child = qq >> clog;
            const int64_t d = child - first_child;
            if (qq >= 0 && qq < n && d >= 0 && d < kMaxFan && (qq >> plog) == region) {
                q[j] = (int)qq;
                dg[j] = (int)d;
                rk[j] = atomicAdd(&s_hist[d], 1u);
            } else {
                atomicOr(err, 1u);
            }
        }
    }
    __syncthreads();
    // Exclusive scan of the 256 counts (one wave) and one global claim per
    // non-empty bucket.
    if (tid < kWave) {
        unsigned c[kMaxFan / kWave], tot = 0;
#pragma unroll
        for (int i = 0; i < kMaxFan / kWave; i++) { c[i] = s_hist[tid * (kMaxFan / kWave) + i]; tot += c[i]; }
        unsigned inc = wave_inclusive_scan(tot, AddU());
        unsigned run = inc - tot;
#pragma unroll
        for (int i = 0; i < kMaxFan / kWave; i++) {
            const int d = tid * (kMaxFan / kWave) + i;
            s_start[d] = run;
            run += c[i];
            if (tile_ofs) s_gofs[d] = tile_ofs[(int64_t)d * gridDim.x + bt];
            else s_gofs[d] = c[i] ? atomicAdd(cnt + first_child + d, c[i]) : 0u;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPartItems; j++) {
        if (dg[j] >= 0) {
            const unsigned at = s_start[dg[j]] + rk[j];
            s_idx[at] = q[j];
            s_val[at] = v[j];
            s_dig[at] = (unsigned char)dg[j];
        }
    }
    __syncthreads();
    const int64_t total = (int64_t)s_start[kMaxFan - 1] + s_hist[kMaxFan - 1];
    const int64_t csize = (int64_t)1 << clog;
    for (int64_t i = tid; i < total; i += kPartThreads) {
        const int d = s_dig[i];
        const int64_t child = first_child + d;
        const int64_t slot = (int64_t)s_gofs[d] + (i - s_start[d]);
        const int64_t cend = (child + 1) * csize < n ? (child + 1) * csize : n;
        const int64_t pos = child * csize + slot;
        if (pos < cend) {
            out.store(pos, s_idx[i], s_val[i]);
        } else {
            atomicOr(err, 1u);   // more pairs than the block holds: not a permutation
        }
    }
}

// Per-tile child counts of the first partition pass (a single region), as
// a 256 x tiles matrix: the column scan of it gives every tile its slot in
// every child without atomics, which matter when thousands of tiles share
// the first pass's few counters.
__global__ __launch_bounds__(kPartThreads) void k_part_hist(IndexSource src, int64_t n, int clog,
                                                            unsigned *__restrict__ hist) {
    __shared__ unsigned s_hist[kMaxFan];
    const int tid = threadIdx.x;
    for (int i = tid; i < kMaxFan; i += kPartThreads) s_hist[i] = 0;
    __syncthreads();
    const int64_t bt = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t lo = bt * kPartTile;
    int64_t qv[kPartItems];
#pragma unroll
    for (int j = 0; j < kPartItems; j++) {
        const int64_t e = lo + (int64_t)j * kPartThreads + tid;
        qv[j] = e < n ? src.at(e) : -1;
    }
#pragma unroll
    for (int j = 0; j < kPartItems; j++) {
        const int64_t qq = qv[j];
        const int64_t d = qq >> clog;
        if (qq >= 0 && qq < n && d < kMaxFan) atomicAdd(&s_hist[d], 1u);
    }
    __syncthreads();
    for (int d = tid; d < kMaxFan; d += kPartThreads) hist[(int64_t)d * gridDim.x + bt] = s_hist[d];
}

// Block scatter: block c owns destination [c*kBlk, ...).  Its pairs are the
// first cnt[c] entries of the same range of (idx, val).
template <class T, bool DIRECT>
__global__ __launch_bounds__(kPartThreads) void k_block_scatter(IndexSource src, Pairs<T> in,
                                                                const T *__restrict__ in_val, T *__restrict__ d,
                                                                int64_t n, const unsigned *__restrict__ cnt,
                                                                unsigned *__restrict__ err) {
    __shared__ T s_out[kBlk];
    __shared__ unsigned s_mask[kBlk / 32];
    const int tid = threadIdx.x;
    const int64_t blk = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t base = blk * kBlk;
    const int64_t end = base + kBlk < n ? base + kBlk : n;
    for (int i = tid; i < kBlk / 32; i += kPartThreads) s_mask[i] = 0;
    __syncthreads();
    int64_t count = end - base;
    if (!DIRECT && (int64_t)cnt[blk] < count) count = cnt[blk];
    for (int64_t i = tid; i < count; i += kPartThreads) {
        int64_t q;
        T v;
        if constexpr (DIRECT) {
            q = src.at(base + i);
            v = src.value(in_val, base + i);
        } else {
            int qi;
            in.load(base + i, qi, v);
            q = qi;
        }
        const int64_t o = q - base;
        if (o >= 0 && o < end - base) {
            s_out[o] = v;
            atomicOr(&s_mask[o >> 5], 1u << (o & 31));
        } else {
            atomicOr(err, 1u);
        }
    }
    __syncthreads();
    for (int64_t o = tid; o < end - base; o += kPartThreads)
        if ((s_mask[o >> 5] >> (o & 31)) & 1u) d[base + o] = s_out[o];
}

static int64_t up256(int64_t b) { return (b + 255) & ~(int64_t)255; }

// Scratch: two (idx, val) buffers of n pairs and one counter array per
// partition pass (at most 3 passes for n < 2^31).
int64_t partition_permute_scratch_bytes(int64_t n, int width) {
    const int64_t pair = up256(4 * n) + up256((int64_t)width * n);
    const int64_t counters = 3 * up256(4 * ((n >> kBlkLog) + 2 + kMaxFan));
    const int64_t hist = up256(4 * (int64_t)kMaxFan * ((n + kPartTile - 1) / kPartTile + 1));
    return 2 * pair + counters + hist + 256;
}

template <class T>
static void run_partition(void *d, const void *s, IndexSource src, int64_t n, void *scratch) {
    unsigned *err = device_error_word();
    // Region sizes per pass: from ceil-pow2(n) down to kBlk, at most 8 bits
    // per pass.
    int nlog = 0;
    while (((int64_t)1 << nlog) < n) nlog++;
    if (nlog <= kBlkLog) {
        k_block_scatter<T, true><<<1, kPartThreads, 0, g_stream>>>(src, Pairs<T>{}, (const T *)s, (T *)d, n, nullptr,
                                                                   err);
        LAUNCHCHECK();
        return;
    }
    if (!scratch) usage_error("permute called without scratch");
    auto up = up256;
    char *b = (char *)scratch;
    Pairs<T> buf[2];
    for (int k = 0; k < 2; k++) {
        if constexpr (sizeof(T) == 4) {
            buf[k].p = (uint2 *)b;
            b += up(8 * n);
        } else {
            buf[k].idx = (int *)b;
            b += up(4 * n);
            buf[k].val = (T *)b;
            b += up((int64_t)sizeof(T) * n);
        }
    }
    const int levels = (nlog - kBlkLog + 7) / 8;
    int plog = nlog;
    int cur = -1;
    unsigned *cnt = nullptr, *prev = nullptr;
    for (int l = 0; l < levels; l++) {
        const int bits = (plog - kBlkLog) - 8 * (levels - 1 - l);   // the first pass takes the odd bits
        const int clog = plog - bits;
        const int64_t nchild = ((n - 1) >> clog) + 1;
        prev = cnt;
        cnt = (unsigned *)b;                       // >= 256 entries: the column scan writes 256 totals
        const int64_t ncnt = (nchild > kMaxFan ? nchild : kMaxFan) + 1;
        b += up(4 * ncnt);
        fill_u32(cnt, 0, ncnt);
        const int64_t regions = ((n - 1) >> plog) + 1;
        const int64_t rsize = (int64_t)1 << plog;
        const int64_t tpr = (rsize < n ? rsize + kPartTile - 1 : n + kPartTile - 1) / kPartTile;
        const int64_t grid = regions * tpr;
        const int out = cur == 0 ? 1 : 0;
        if (cur < 0) {
            // first pass: one region, many tiles -> slots from a histogram + column scan
            unsigned *hist = (unsigned *)b;
            b += up(4 * (int64_t)kMaxFan * grid);
            k_part_hist<<<(unsigned)grid, kPartThreads, 0, g_stream>>>(src, n, clog, hist);
            LAUNCHCHECK();
            k_col_scan<<<kMaxFan, kColScanThreads, 0, g_stream>>>(hist, grid, cnt);
            LAUNCHCHECK();
            k_partition<T, true><<<(unsigned)grid, kPartThreads, 0, g_stream>>>(
                src, Pairs<T>{}, (const T *)s, buf[out], n, plog, clog, tpr, nullptr, cnt, hist, err);
        } else {
            k_partition<T, false><<<(unsigned)grid, kPartThreads, 0, g_stream>>>(
                src, buf[cur], nullptr, buf[out], n, plog, clog, tpr, prev, cnt, nullptr, err);
        }
        LAUNCHCHECK();
        cur = out;
        plog = clog;
    }
    const int64_t blocks = ((n - 1) >> kBlkLog) + 1;
    k_block_scatter<T, false><<<(unsigned)blocks, kPartThreads, 0, g_stream>>>(src, buf[cur], nullptr, (T *)d, n,
                                                                              cnt, err);
    LAUNCHCHECK();
}

template <class T>
void partition_permute(void *d, const void *s, const int *p, const void *sd, int64_t n, int64_t m, void *scratch) {
    if (n <= 0) return;
    IndexSource src{p, SegdView(sd, n, sd ? m : 0), sd != nullptr && m > 1, 0, SegdView()};
    run_partition<T>(d, s, src, n, scratch);
}

void partition_invert(int *rank, const int *order, int64_t n, const void *sd, int64_t m, void *scratch) {
    if (n <= 0) return;
    const bool rel = sd != nullptr && m > 1;
    IndexSource src{order, SegdView(), false, rel ? 2 : 1, rel ? SegdView(sd, n, m) : SegdView()};
    run_partition<int>(rank, nullptr, src, n, scratch);
}

template void partition_permute<int>(void *, const void *, const int *, const void *, int64_t, int64_t, void *);
template void partition_permute<double>(void *, const void *, const int *, const void *, int64_t, int64_t, void *);

}  // namespace cvlhip
