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
//                     region into <= 256 children; two passes (8 bits, then
//                     7) take 2^28 elements to 8192-element blocks.
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
#include "lookback.hpp"
#include "../../include/cvl.h"

namespace cvlhip {

// Final destination blocks are LDS-resident in the block scatter: 2^kBlkLog4
// four-byte or 2^kBlkLog8 eight-byte values per workgroup.
#ifndef CVLHIP_BLK_LOG4
#define CVLHIP_BLK_LOG4 13
#endif
#ifndef CVLHIP_BLK_LOG8
#define CVLHIP_BLK_LOG8 13
#endif
constexpr int kBlkLog4 = CVLHIP_BLK_LOG4, kBlkLog8 = CVLHIP_BLK_LOG8;
constexpr int kBlkLogMin = kBlkLog4 < kBlkLog8 ? kBlkLog4 : kBlkLog8;
template <class T>
constexpr int blk_log() { return sizeof(T) == 4 ? kBlkLog4 : kBlkLog8; }
#ifndef CVLHIP_PART_ITEMS
#define CVLHIP_PART_ITEMS 8
#endif
#ifndef CVLHIP_PART_THREADS
#define CVLHIP_PART_THREADS 1024
#endif
// 1: the first pass chains each tile's slots in the children from the tiles
// before it (look-back) instead of a histogram pass over the indices and a
// column scan.
#ifndef CVLHIP_PART_LOOKBACK
#define CVLHIP_PART_LOOKBACK 0
#endif
constexpr bool kPartLookback = CVLHIP_PART_LOOKBACK != 0;
// 1: the first pass claims its slots like the later passes, one atomic per
// (tile, child), instead of a histogram pass over the indices and a
// column scan.
#ifndef CVLHIP_PART_ATOMIC_FIRST
#define CVLHIP_PART_ATOMIC_FIRST 0
#endif
constexpr bool kPartAtomicFirst = CVLHIP_PART_ATOMIC_FIRST != 0 && !kPartLookback;
constexpr int kPartThreads = CVLHIP_PART_THREADS;
constexpr int kPartItems = CVLHIP_PART_ITEMS;
constexpr int kPartTile = kPartThreads * kPartItems;   // pairs per tile
// Item j of a thread is element lo + ((j / kPartV) * kPartThreads + tid) *
// kPartV + j % kPartV of its tile: each lane owns runs of kPartV consecutive
// elements, read with 16-byte loads where the input is aligned.  (Order
// within a bucket is free: the indices of a permutation are distinct.)
constexpr int kPartV = 4;
static_assert(kPartItems % kPartV == 0, "items per thread come in runs of kPartV");
constexpr int kMaxFan = 256;                   // children per region per pass
static_assert(kPartThreads >= kMaxFan, "every child has a thread for its slot");
static_assert(kMaxFan == kTileScanCols, "first-pass histogram rows are scanned 256 wide");
static_assert(kBlkLog4 <= 16 && kBlkLog8 <= 16, "offsets within a final block fit the 16 bits the last pass keeps");

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
    bool vec;         // unsegmented, vmode 0 or 1, p (and s) 16-byte aligned: runs load as vectors
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
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

// One vector element reinterpreted as T.  The element is passed by value,
// so the cast sees a scalar: __builtin_bit_cast applied directly to an
// ext_vector subscript (w[c]) reads element 0 whatever c is with this
// compiler.
template <class T, class U>
__device__ __forceinline__ T elem_as(U u) {
    return __builtin_bit_cast(T, u);
}

// Four consecutive values from a 16-byte-aligned address.
template <class T>
__device__ __forceinline__ void load4(const T *__restrict__ p, T (&x)[4]) {
    if constexpr (sizeof(T) == 4) {
        const u32x4 w = *(const u32x4 *)p;
#pragma unroll
        for (int c = 0; c < 4; c++) x[c] = elem_as<T>(w[c]);
    } else {
        const u64x2 a = *(const u64x2 *)p, b = *(const u64x2 *)(p + 2);
        x[0] = elem_as<T>(a[0]);
        x[1] = elem_as<T>(a[1]);
        x[2] = elem_as<T>(b[0]);
        x[3] = elem_as<T>(b[1]);
    }
}

template <class T, bool PACKED = sizeof(T) == 4>
struct Pairs {
    int *idx;
    T *val;
    static constexpr bool LOCAL = false;
    __device__ __forceinline__ void load(int64_t e, int &q, T &v) const {
        q = idx[e];
        v = val[e];
    }
    // pairs e0 .. e0+3, e0 a multiple of 4 (the buffers are 256-byte aligned)
    __device__ __forceinline__ void load4(int64_t e0, int (&q)[4], T (&v)[4]) const {
        cvlhip::load4(idx + e0, q);
        cvlhip::load4(val + e0, v);
    }
    __device__ __forceinline__ void store(int64_t e, int q, T v) const {
        idx[e] = q;
        val[e] = v;
    }
};
template <class T>
struct Pairs<T, true> {
    uint2 *p;
    static constexpr bool LOCAL = false;
    __device__ __forceinline__ void load(int64_t e, int &q, T &v) const {
        const uint2 w = p[e];
        q = (int)w.x;
        v = __builtin_bit_cast(T, w.y);
    }
    __device__ __forceinline__ void load4(int64_t e0, int (&q)[4], T (&v)[4]) const {
        const u32x4 a = *(const u32x4 *)(p + e0), b = *(const u32x4 *)(p + e0 + 2);
        q[0] = (int)a[0]; v[0] = elem_as<T>(a[1]);
        q[1] = (int)a[2]; v[1] = elem_as<T>(a[3]);
        q[2] = (int)b[0]; v[2] = elem_as<T>(b[1]);
        q[3] = (int)b[2]; v[3] = elem_as<T>(b[3]);
    }
    __device__ __forceinline__ void store(int64_t e, int q, T v) const {
        p[e] = make_uint2((unsigned)q, __builtin_bit_cast(unsigned, v));
    }
};

// Output of the last partition pass, read by the block scatter: the value
// and the destination index's low 16 bits, whose low blk_log<T>() bits are
// the offset within the destination block (a block starts at a multiple of
// its size).  Six bytes per four-byte value instead of eight.
template <class T>
struct Local16 {
    unsigned short *idx;
    T *val;
    static constexpr bool LOCAL = true;
    __device__ __forceinline__ void load(int64_t e, int &q, T &v) const {
        q = idx[e];
        v = val[e];
    }
    // entries e0 .. e0+3, e0 a multiple of 4 (the buffers are 256-byte aligned)
    __device__ __forceinline__ void load4(int64_t e0, int (&q)[4], T (&v)[4]) const {
        typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
        const u16x4 w = *(const u16x4 *)(idx + e0);
#pragma unroll
        for (int c = 0; c < 4; c++) q[c] = (int)(unsigned)w[c];
        cvlhip::load4(val + e0, v);
    }
    __device__ __forceinline__ void store(int64_t e, int q, T v) const {
        idx[e] = (unsigned short)q;
        val[e] = v;
    }
};

// One partition pass.  Region r = [r*psize, min((r+1)*psize, n)) of the
// input is split into children of csize elements (csize divides psize,
// psize/csize <= 256).  Block b handles tile (b % tiles_per_region) of
// region (b / tiles_per_region).  Child c's pairs land in
// [c*csize, c*csize + cnt[c]) of the output.
// A first pass takes each tile's slot in each child from tile_ofs (the
// column scan of per-tile histograms) or, with `lb`, from a look-back over
// the tiles before it: tiles are then claimed in order from lb_ctr, every
// tile publishes its counts and adds up the earlier ones child by child
// (a tile waits only on tiles already claimed, so every wait ends), and the
// last tile writes the children's totals to cnt.
template <class T, bool FIRST, class Out = Pairs<T>>
__global__ __launch_bounds__(kPartThreads) void k_partition(IndexSource src, Pairs<T> in,
                                                            const T *__restrict__ in_val, Out out,
                                                            int64_t n, int plog, int clog,
                                                            int64_t tiles_per_region,
                                                            const unsigned *__restrict__ prev_cnt,
                                                            unsigned *__restrict__ cnt,
                                                            const unsigned *__restrict__ tile_ofs,
                                                            unsigned *__restrict__ err,
                                                            uint32_t *__restrict__ lb = nullptr,
                                                            unsigned *__restrict__ lb_ctr = nullptr) {
    __shared__ unsigned s_hist[kMaxFan];
    __shared__ unsigned s_start[kMaxFan];
    __shared__ unsigned s_gofs[kMaxFan];
    __shared__ unsigned s_bt;
    // The tile staged in bucket order: four-byte values share one 8-byte
    // word with their index (one LDS access each way, and the bucket is
    // recomputed from the index); wider values keep the index, the value
    // and the bucket apart.
    constexpr bool PACKED = sizeof(T) == 4;
    __shared__ uint2 s_pair[PACKED ? kPartTile : 1];
    __shared__ int s_idx[PACKED ? 1 : kPartTile];
    __shared__ T s_val[PACKED ? 1 : kPartTile];
    __shared__ unsigned char s_dig[PACKED ? 1 : kPartTile];

    const int tid = threadIdx.x;
    constexpr bool LB = FIRST && kPartLookback;    // compiled in only when selected
    for (int i = tid; i < kMaxFan; i += kPartThreads) s_hist[i] = 0;
    if (LB && tid == 0) s_bt = atomicAdd(lb_ctr, 1u);
    __syncthreads();
    const int64_t bt = LB ? (int64_t)s_bt : xcd_tile(blockIdx.x, gridDim.x);
    const int64_t region = bt / tiles_per_region;
    const int64_t chunk = bt % tiles_per_region;
    const int64_t rlo = region << plog;
    int64_t rhi = (rlo + ((int64_t)1 << plog)) < n ? rlo + ((int64_t)1 << plog) : n;
    if (!FIRST && rlo + (int64_t)prev_cnt[region] < rhi) rhi = rlo + prev_cnt[region];   // pairs that landed
    const int64_t lo = rlo + chunk * kPartTile;
    const int64_t first_child = rlo >> clog;

    int q[kPartItems];
    T v[kPartItems];
    unsigned rk[kPartItems];
    int dg[kPartItems];
    int64_t qv[kPartItems];
    bool has[kPartItems];
#pragma unroll
    for (int g = 0; g < kPartItems / kPartV; g++) {
        const int64_t e0 = lo + ((int64_t)g * kPartThreads + tid) * kPartV;
        if (e0 + kPartV <= rhi && (!FIRST || src.vec)) {
            int qi[kPartV];
            T vi[kPartV];
            if constexpr (FIRST) {
                load4(src.p + e0, qi);
                if (src.vmode == 0) {
                    load4(in_val + e0, vi);
                } else {
#pragma unroll
                    for (int c = 0; c < kPartV; c++) vi[c] = (T)(e0 + c);
                }
            } else {
                in.load4(e0, qi, vi);
            }
#pragma unroll
            for (int c = 0; c < kPartV; c++) {
                qv[g * kPartV + c] = qi[c];
                v[g * kPartV + c] = vi[c];
                has[g * kPartV + c] = true;
            }
        } else {
#pragma unroll
            for (int c = 0; c < kPartV; c++) {
                const int j = g * kPartV + c;
                const int64_t e = e0 + c;
                has[j] = e < rhi;
                qv[j] = -1;
                if (has[j]) {
                    if constexpr (FIRST) {
                        qv[j] = src.at(e);
                        v[j] = src.value(in_val, e);
                    } else {
                        int qi;
                        in.load(e, qi, v[j]);
                        qv[j] = qi;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kPartItems; j++) {
        dg[j] = -1;
        if (has[j]) {
            const int64_t qq = qv[j];
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
    // Exclusive scan of the 256 counts (one wave), and the tile's slot in
    // every child: from the column scan (first pass) or one global claim per
    // non-empty bucket, one thread per bucket so the claims are in flight
    // together.
    if (tid < kWave) {
        unsigned c[kMaxFan / kWave], tot = 0;
#pragma unroll
        for (int i = 0; i < kMaxFan / kWave; i++) { c[i] = s_hist[tid * (kMaxFan / kWave) + i]; tot += c[i]; }
        unsigned inc = wave_inclusive_scan(tot, AddU());
        unsigned run = inc - tot;
#pragma unroll
        for (int i = 0; i < kMaxFan / kWave; i++) {
            s_start[tid * (kMaxFan / kWave) + i] = run;
            run += c[i];
        }
    }
    if (!LB && tid < kMaxFan) {
        const int d = tid;
        if (tile_ofs) {
            s_gofs[d] = tile_ofs[bt * kMaxFan + d];
        } else {
            const unsigned c = s_hist[d];
            s_gofs[d] = c ? atomicAdd(cnt + first_child + d, c) : 0u;
        }
    }
    if (LB && tid < kMaxFan) {
        // one thread per child: the counts of the tiles before this one
        const int d = tid;
        const unsigned c = s_hist[d];
        const unsigned excl = chain_exclusive(lb, bt, kMaxFan, d, c, 0, err);
        s_gofs[d] = excl;
        if (bt == (int64_t)gridDim.x - 1) cnt[first_child + d] = excl + c;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPartItems; j++) {
        if (dg[j] >= 0) {
            const unsigned at = s_start[dg[j]] + rk[j];
            if constexpr (PACKED) {
                s_pair[at] = make_uint2((unsigned)q[j], elem_as<unsigned>(v[j]));
            } else {
                s_idx[at] = q[j];
                s_val[at] = v[j];
                s_dig[at] = (unsigned char)dg[j];
            }
        }
    }
    __syncthreads();
    const int64_t total = (int64_t)s_start[kMaxFan - 1] + s_hist[kMaxFan - 1];
    const int64_t csize = (int64_t)1 << clog;
    for (int64_t i = tid; i < total; i += kPartThreads) {
        int qi;
        T vi;
        int d;
        if constexpr (PACKED) {
            const uint2 w = s_pair[i];
            qi = (int)w.x;
            vi = __builtin_bit_cast(T, w.y);
            d = (int)(((int64_t)qi >> clog) - first_child);   // in [0, 256): checked when ranked
        } else {
            qi = s_idx[i];
            vi = s_val[i];
            d = s_dig[i];
        }
        const int64_t child = first_child + d;
        const int64_t slot = (int64_t)s_gofs[d] + (i - s_start[d]);
        const int64_t cend = (child + 1) * csize < n ? (child + 1) * csize : n;
        const int64_t pos = child * csize + slot;
        if (pos < cend) {
            out.store(pos, qi, vi);
        } else {
            atomicOr(err, 1u);   // more pairs than the block holds: not a permutation
        }
    }
}

// Per-tile child counts of the first partition pass (a single region), as
// a tiles x 256 matrix: the column scan of it gives every tile its slot in
// every child without atomics, which matter when thousands of tiles share
// the first pass's few counters.
__global__ __launch_bounds__(kPartThreads) void k_part_hist(IndexSource src, int64_t n, int clog,
                                                            unsigned *__restrict__ hist) {
    __shared__ unsigned s_hist[kMaxFan];
    const int tid = threadIdx.x;
    for (int i = tid; i < kMaxFan; i += kPartThreads) s_hist[i] = 0;
    __syncthreads();
    // Rows are whole 1 KiB stores, so tiles need no XCD grouping here.
    const int64_t bt = blockIdx.x;
    const int64_t lo = bt * kPartTile;
    if (src.vec && lo + kPartTile <= n) {
        // A whole tile of a plain index vector: 16-byte loads, and one
        // unsigned compare rejects both negative and too-large indices (n
        // >> clog <= 256, so every index below n has a digit below 256).
        static_assert(kPartV == 4, "runs of four indices per 16-byte load");
        u32x4 w[kPartItems / kPartV];
#pragma unroll
        for (int g = 0; g < kPartItems / kPartV; g++)
            w[g] = *(const u32x4 *)(src.p + lo + ((int64_t)g * kPartThreads + tid) * kPartV);
#pragma unroll
        for (int g = 0; g < kPartItems / kPartV; g++) {
            const unsigned q[4] = {w[g].x, w[g].y, w[g].z, w[g].w};
#pragma unroll
            for (int c = 0; c < 4; c++)
                if (q[c] < (uint64_t)n) atomicAdd(&s_hist[q[c] >> clog], 1u);
        }
        __syncthreads();
        for (int d = tid; d < kMaxFan; d += kPartThreads) hist[bt * kMaxFan + d] = s_hist[d];
        return;
    }
    int64_t qv[kPartItems];
#pragma unroll
    for (int g = 0; g < kPartItems / kPartV; g++) {
        const int64_t e0 = lo + ((int64_t)g * kPartThreads + tid) * kPartV;
        if (src.vec && e0 + kPartV <= n) {
            int qi[kPartV];
            load4(src.p + e0, qi);
#pragma unroll
            for (int c = 0; c < kPartV; c++) qv[g * kPartV + c] = qi[c];
        } else {
#pragma unroll
            for (int c = 0; c < kPartV; c++) qv[g * kPartV + c] = e0 + c < n ? src.at(e0 + c) : -1;
        }
    }
#pragma unroll
    for (int j = 0; j < kPartItems; j++) {
        const int64_t qq = qv[j];
        const int64_t d = qq >> clog;
        if (qq >= 0 && qq < n && d < kMaxFan) atomicAdd(&s_hist[d], 1u);
    }
    __syncthreads();
    for (int d = tid; d < kMaxFan; d += kPartThreads) hist[bt * kMaxFan + d] = s_hist[d];
}

// Block scatter: block c owns destination [c*kBlk, ...).  Its pairs are the
// first cnt[c] entries of the same range of (idx, val).
template <class T, bool DIRECT, class In = Pairs<T>>
__global__ __launch_bounds__(kPartThreads) void k_block_scatter(IndexSource src, In in,
                                                                const T *__restrict__ in_val, T *__restrict__ d,
                                                                int64_t n, const unsigned *__restrict__ cnt,
                                                                unsigned *__restrict__ err) {
    constexpr int64_t kBlk = (int64_t)1 << blk_log<T>();
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
    auto put = [&](int64_t q, T v) {
        const int64_t o = In::LOCAL ? (q & (kBlk - 1)) : q - base;
        if (o >= 0 && o < end - base) {
            s_out[o] = v;
            atomicOr(&s_mask[o >> 5], 1u << (o & 31));
        } else {
            atomicOr(err, 1u);
        }
    };
    if constexpr (DIRECT) {
        for (int64_t i = tid; i < count; i += kPartThreads) put(src.at(base + i), src.value(in_val, base + i));
    } else {
        // runs of 4 pairs per lane (base is a multiple of 4)
        for (int64_t i0 = (int64_t)tid * 4; i0 < count; i0 += (int64_t)kPartThreads * 4) {
            int qi[4];
            T vi[4];
            if (i0 + 4 <= count) {
                in.load4(base + i0, qi, vi);
#pragma unroll
                for (int c = 0; c < 4; c++) put(qi[c], vi[c]);
            } else {
                for (int64_t i = i0; i < count; i++) {
                    in.load(base + i, qi[0], vi[0]);
                    put(qi[0], vi[0]);
                }
            }
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
    const int64_t counters = 3 * up256(4 * ((n >> kBlkLogMin) + 2 + kMaxFan));
    const int64_t tiles = (n + kPartTile - 1) / kPartTile + 1;
    const int64_t hist = up256(4 * (int64_t)kMaxFan * tiles) + up256(8 * tile_scan_status_words(tiles));
    return 2 * pair + counters + hist + 256;
}

// The scratch of one partition run: the two pair buffers (the last pass
// writes Local16 into the space of the buffer it would otherwise fill:
// 2 + sizeof(T) <= the pair's bytes per element), one child-counter array
// per level, then the first pass's column-scan status and histogram.
// clear() zeroes the counters and the status (one launch).
template <class T>
struct PartLayout {
    int nlog = 0, levels = 0;
    Pairs<T> buf[2];
    Local16<T> local[2];
    unsigned *cnts[4] = {};
    uint64_t *status0 = nullptr;
    unsigned *hist0 = nullptr;
    int64_t grid0 = 0;            // tiles of the first pass (one region)
    // Region sizes per pass: from ceil-pow2(n) down to the final block, at
    // most 8 bits per pass; full 8-bit passes first, the odd bits last (the
    // first pass takes its slots from the histogram, later passes claim
    // them with one atomic per (tile, child), so the narrower fan belongs
    // there).
    int bits(int l, int plog) const { return l + 1 < levels ? 8 : plog - blk_log<T>(); }
    char *clear_from = nullptr, *clear_to = nullptr;
    void clear() const {
        if (clear_to > clear_from) fill_u32((unsigned *)clear_from, 0, (int64_t)(clear_to - clear_from) / 4);
    }
    PartLayout(int64_t n, void *scratch) {
        while (((int64_t)1 << nlog) < n) nlog++;
        if (nlog <= blk_log<T>()) return;   // one block: no scratch
        if (!scratch) usage_error("permute called without scratch");
        auto up = up256;
        char *b = (char *)scratch;
        for (int k = 0; k < 2; k++) {
            char *const b0 = b;
            if constexpr (sizeof(T) == 4) {
                buf[k].p = (uint2 *)b;
                b += up(8 * n);
            } else {
                buf[k].idx = (int *)b;
                b += up(4 * n);
                buf[k].val = (T *)b;
                b += up((int64_t)sizeof(T) * n);
            }
            local[k].idx = (unsigned short *)b0;
            local[k].val = (T *)(b0 + up(2 * n));
        }
        levels = (nlog - blk_log<T>() + 7) / 8;
        // per level: its child counters (>= 256 entries: the column scan
        // writes 256 totals)
        int pl = nlog;
        for (int l = 0; l < levels; l++) {
            const int cl = pl - bits(l, pl);
            const int64_t nchild = ((n - 1) >> cl) + 1;
            cnts[l] = (unsigned *)b;
            b += up(4 * ((nchild > kMaxFan ? nchild : kMaxFan) + 1));
            pl = cl;
        }
        grid0 = (n + kPartTile - 1) / kPartTile;
        // the column scan's status, or the look-back's tile counter; then the
        // histogram, or the look-back words (cleared with the counters)
        status0 = (uint64_t *)b;
        b += up(8 * tile_scan_status_words(grid0));
        hist0 = (unsigned *)b;
        if (kPartLookback) b += up(4 * (int64_t)kMaxFan * grid0);
        clear_from = (char *)cnts[0];
        clear_to = b;
    }
};

// Levels l0 .. levels-1 after the first (their input: buffer `cur`, its
// children counted in cnts[l0 - 1]), then the block scatter into d.
template <class T>
static void partition_later_levels(const PartLayout<T> &L, IndexSource src, void *d, int64_t n, int l0, int cur,
                                   unsigned *err) {
    int plog = L.nlog;
    for (int l = 0; l < l0; l++) plog -= L.bits(l, plog);
    unsigned *cnt = L.cnts[l0 - 1];
    for (int l = l0; l < L.levels; l++) {
        const int clog = plog - L.bits(l, plog);
        unsigned *const prev = cnt;
        cnt = L.cnts[l];
        const int64_t regions = ((n - 1) >> plog) + 1;
        const int64_t rsize = (int64_t)1 << plog;
        const int64_t tpr = (rsize < n ? rsize + kPartTile - 1 : n + kPartTile - 1) / kPartTile;
        const int64_t grid = regions * tpr;
        const int out = cur == 0 ? 1 : 0;
        if (l + 1 == L.levels)        // children are the final blocks: write Local16
            k_partition<T, false, Local16<T>><<<(unsigned)grid, kPartThreads, 0, g_stream>>>(
                src, L.buf[cur], nullptr, L.local[out], n, plog, clog, tpr, prev, cnt, nullptr, err);
        else
            k_partition<T, false><<<(unsigned)grid, kPartThreads, 0, g_stream>>>(
                src, L.buf[cur], nullptr, L.buf[out], n, plog, clog, tpr, prev, cnt, nullptr, err);
        LAUNCHCHECK();
        cur = out;
        plog = clog;
    }
    const int64_t blocks = ((n - 1) >> blk_log<T>()) + 1;
    if (L.levels == l0) usage_error("permute: no level after the first");
    k_block_scatter<T, false, Local16<T>><<<(unsigned)blocks, kPartThreads, 0, g_stream>>>(
        src, L.local[cur], nullptr, (T *)d, n, cnt, err);
    LAUNCHCHECK();
}

template <class T>
static void run_partition(void *d, const void *s, IndexSource src, int64_t n, void *scratch) {
    unsigned *err = device_error_word();
    const PartLayout<T> L(n, scratch);
    L.clear();
    if (L.levels == 0) {
        k_block_scatter<T, true><<<1, kPartThreads, 0, g_stream>>>(src, Pairs<T>{}, (const T *)s, (T *)d, n, nullptr,
                                                                   err);
        LAUNCHCHECK();
        return;
    }
    // first pass: one region, many tiles -> slots from a histogram + column
    // scan (or the look-back)
    const int plog = L.nlog;
    const int clog = plog - L.bits(0, plog);
    const int64_t tpr = L.grid0;
    unsigned *cnt = L.cnts[0];
    unsigned *hist = (kPartLookback || kPartAtomicFirst) ? nullptr : L.hist0;
    uint32_t *lb = kPartLookback ? L.hist0 : nullptr;
    unsigned *lb_ctr = kPartLookback ? (unsigned *)L.status0 : nullptr;
    if (!kPartLookback && !kPartAtomicFirst) {
        k_part_hist<<<(unsigned)L.grid0, kPartThreads, 0, g_stream>>>(src, n, clog, hist);
        LAUNCHCHECK();
        tile_scan(hist, L.grid0, cnt, L.status0, true);
    }
    if (L.levels == 1) {
        k_partition<T, true, Local16<T>><<<(unsigned)L.grid0, kPartThreads, 0, g_stream>>>(
            src, Pairs<T>{}, (const T *)s, L.local[0], n, plog, clog, tpr, nullptr, cnt, hist, err, lb, lb_ctr);
        LAUNCHCHECK();
        const int64_t blocks = ((n - 1) >> blk_log<T>()) + 1;
        k_block_scatter<T, false, Local16<T>><<<(unsigned)blocks, kPartThreads, 0, g_stream>>>(
            src, L.local[0], nullptr, (T *)d, n, cnt, err);
        LAUNCHCHECK();
        return;
    }
    k_partition<T, true><<<(unsigned)L.grid0, kPartThreads, 0, g_stream>>>(
        src, Pairs<T>{}, (const T *)s, L.buf[0], n, plog, clog, tpr, nullptr, cnt, hist, err, lb, lb_ctr);
    LAUNCHCHECK();
    partition_later_levels<T>(L, src, d, n, 1, 0, err);
}

template <class T>
void partition_permute(void *d, const void *s, const int *p, const void *sd, int64_t n, int64_t m, void *scratch) {
    if (n <= 0) return;
    const bool segmented = sd != nullptr && m > 1;
    const bool vec = !segmented && ((uintptr_t)p & 15) == 0 && ((uintptr_t)s & 15) == 0;
    IndexSource src{p, SegdView(sd, n, sd ? m : 0), segmented, 0, SegdView(), vec};
    run_partition<T>(d, s, src, n, scratch);
}

// The inverse of a sort order in two steps: the caller runs the first
// partition pass (the rank's last pass, which knows each element's final
// position as it places it) and partition_invert_finish the rest.
InvertFirstPass partition_invert_begin(int64_t n, void *scratch) {
    InvertFirstPass f{};
    if (n <= 0) return f;
    const PartLayout<int> L(n, scratch);
    if (L.levels < 2) return f;          // one partition pass or none: partition_invert
    L.clear();
    f.pairs = L.buf[0].p;
    f.cnt = L.cnts[0];
    f.clog = L.nlog - L.bits(0, L.nlog);
    return f;
}

void partition_invert_finish(int *rank, int64_t n, void *scratch) {
    const PartLayout<int> L(n, scratch);   // the layout of partition_invert_begin, counters kept
    if (L.levels < 2) usage_error("partition_invert_finish: no split invert for %lld elements", (long long)n);
    // later passes read only the pair buffers; the source is not consulted
    const IndexSource src{nullptr, SegdView(), false, 1, SegdView(), true};
    partition_later_levels<int>(L, src, rank, n, 1, 0, device_error_word());
}

void partition_invert(int *rank, const int *order, int64_t n, const void *sd, int64_t m, void *scratch) {
    if (n <= 0) return;
    const bool rel = sd != nullptr && m > 1;
    const bool vec = !rel && ((uintptr_t)order & 15) == 0;
    IndexSource src{order, SegdView(), false, rel ? 2 : 1, rel ? SegdView(sd, n, m) : SegdView(), vec};
    run_partition<int>(rank, nullptr, src, n, scratch);
}

template void partition_permute<int>(void *, const void *, const int *, const void *, int64_t, int64_t, void *);
template void partition_permute<double>(void *, const void *, const int *, const void *, int64_t, int64_t, void *);

}  // namespace cvlhip
