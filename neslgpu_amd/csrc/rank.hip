// rank.hip -- stable radix rank: rank[i] = position of s[i] in the stable
// sort of s (ref nesl/cvl/serial/rank.c:40-362).
//
// Same digit sequence as the reference: 8-bit digits, least significant
// first, over the transformed 32-bit key (int: key ^ 0x80000000 up,
// key ^ 0x7fffffff down) or the two transformed 32-bit fields of a double
// (low word first), then, for segmented ranks, over the segment number; the
// rank is the final position minus the segment's offset.  A stable LSD sort
// is unique, so any pass that sees a single digit value can be skipped
// without changing the result: one histogram pass reads all keys first and
// the host launches only the passes that reorder something.
//
// Each reordering pass is a histogram kernel (per-tile digit counts), a
// column scan of those counts (each tile's offset in each digit's bucket)
// and a scatter kernel -- or, with CVLHIP_RANK_CHAIN, the scatter kernel
// alone, its tiles chaining the digit offsets with a look-back (tiles
// claimed in order, one chained word per tile and digit): a tile of 4096
// elements ranks its digits stably with a wave-level multisplit (8 ballots
// give each lane the lanes holding the same digit), stages the tile in LDS in
// digit order and writes each digit's run contiguously (keys, then
// indices).  The last pass leaves the sorted order of indices; the rank is
// its inverse, rank[order[p]] = p (minus the segment offset), computed by
// the coalesced permutation engine of partition.hip.
#include "cvl_hip.hpp"
#include "segd.hpp"
#include "lookback.hpp"
#include "partition.hpp"
#include "colscan.hpp"
#include "../../include/cvl.h"

namespace cvlhip {

constexpr int kRadix = 256;
static_assert(kRadix == kTileScanCols, "per-tile digit counts are scanned 256 wide");
#ifndef CVLHIP_RANK_THREADS
#define CVLHIP_RANK_THREADS 512     // 4096-key tiles: 2.03 -> 1.97 ms at 2^26 against 8192
#endif
#ifndef CVLHIP_RANK_ITEMS
#define CVLHIP_RANK_ITEMS 8
#endif
// 1: each pass chains its tiles' digit offsets with a look-back inside the
// scatter kernel (tiles claimed in order) instead of a per-tile histogram
// kernel and a column scan before it.
#ifndef CVLHIP_RANK_CHAIN
#define CVLHIP_RANK_CHAIN 0
#endif
constexpr bool kRankChain = CVLHIP_RANK_CHAIN != 0;
constexpr int kRankThreads = CVLHIP_RANK_THREADS;        // waves of 64
static_assert(kRankThreads >= 256, "one thread per digit");
constexpr int kRankItems = CVLHIP_RANK_ITEMS;            // per lane
constexpr int kRankTile = kRankThreads * kRankItems;     // keys per tile
constexpr int kWaveSpan = kWave * kRankItems;            // consecutive keys per wave

// ---------------------------------------------------------------- key transforms
// KeyT: uint32_t (int ranks) or uint64_t (double ranks, low field in the low
// 32 bits).
__device__ __forceinline__ uint32_t int_key(int s, bool up) {
    return (uint32_t)s ^ (up ? 0x80000000u : 0x7fffffffu);
}
__device__ __forceinline__ uint64_t dbl_key(double s, bool up) {
    uint64_t b = __builtin_bit_cast(uint64_t, s);
    uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32);
    bool neg = (b >> 63) != 0;
    uint32_t flo = neg ? ~lo : lo;
    uint32_t fhi = neg ? ~hi : (hi ^ 0x80000000u);
    if (!up) { flo = ~flo; fhi = ~fhi; }
    return ((uint64_t)fhi << 32) | flo;
}

// ---------------------------------------------------------------- histogram
// hist[p][d]: number of elements whose digit in pass p is d, for the key
// passes (from the keys) -- segment passes are counted by k_seg_hist.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Keys are read 16 bytes per lane when the input is 16-byte aligned
// (`aligned`); the counts do not depend on which lane sees which key.
template <bool DBL>
__global__ __launch_bounds__(kBlock) void k_key_hist(const void *__restrict__ s, int64_t n, bool up, bool aligned,
                                                     unsigned *__restrict__ hist) {
    constexpr int P = DBL ? 8 : 4;
    constexpr int V = DBL ? 2 : 4;               // keys per 16 bytes
    __shared__ unsigned sh[P][kRadix];
    for (int i = threadIdx.x; i < P * kRadix; i += kBlock) (&sh[0][0])[i] = 0;
    __syncthreads();
    auto count = [&](uint64_t key) {
#pragma unroll
        for (int p = 0; p < P; p++) atomicAdd(&sh[p][(key >> (8 * p)) & 255], 1u);
    };
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    const int64_t nv = aligned ? n / V : 0;
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < nv; k += stride) {
        const u32x4 w = ((const u32x4 *)s)[k];
        if constexpr (DBL) {
            count(dbl_key(__builtin_bit_cast(double, ((uint64_t)w.y << 32) | w.x), up));
            count(dbl_key(__builtin_bit_cast(double, ((uint64_t)w.w << 32) | w.z), up));
        } else {
            count(int_key((int)w.x, up));
            count(int_key((int)w.y, up));
            count(int_key((int)w.z, up));
            count(int_key((int)w.w, up));
        }
    }
    for (int64_t k = nv * V + (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += stride)
        count(DBL ? dbl_key(((const double *)s)[k], up) : int_key(((const int *)s)[k], up));
    __syncthreads();
    for (int i = threadIdx.x; i < P * kRadix; i += kBlock)
        if ((&sh[0][0])[i]) atomicAdd(hist + i, (&sh[0][0])[i]);
}

// Which passes run.  A pass whose digit takes one value over all keys keeps
// the order and is skipped; the digit takes one value iff none of its bits
// is 1 in one key and 0 in another, i.e. shares no bit with (OR of the keys)
// AND (OR of the complemented keys).  The keys' OR words come from the
// first pass's histogram kernel; those of the segment numbers (segmented
// ranks: the numbers of the non-empty segments) from k_seg_or, as `words`
// [0] (OR of the numbers) and [1] (OR of their complements), one per
// 256-byte line, in kOrCopies copies (workgroup b adds to copy
// b % kOrCopies, so fewer atomics queue on each word).
constexpr int kOrStride = 64;
constexpr int kOrWords = 2;
constexpr int kOrCopies = 8;
struct OrU {
    using T = uint32_t;
    __host__ __device__ static constexpr uint32_t neutral() { return 0u; }
    __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return a | b; }
};

// The workgroup's OR of v[0..NW) into words[first..first+NW): one atomic
// per word and workgroup.
template <int NW>
__device__ __forceinline__ void or_publish(const uint32_t (&v)[NW], unsigned *__restrict__ words, int first) {
    __shared__ uint32_t s_or[kBlock / kWave][NW];
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
    for (int k = 0; k < NW; k++) {
        const uint32_t r = wave_reduce(v[k], OrU());
        if (lane == 0) s_or[w][k] = r;
    }
    __syncthreads();
    if (threadIdx.x < NW) {
        uint32_t r = 0;
#pragma unroll
        for (int i = 0; i < kBlock / kWave; i++) r |= s_or[i][threadIdx.x];
        if (r) atomicOr(words + ((blockIdx.x % kOrCopies) * kOrWords + first + threadIdx.x) * kOrStride, r);
    }
}

// The runs: the per-tile words of the first pass's keys (k_rank_hist with
// ORW: bits that differ from key 0, two words per tile) folded with the
// segment words of k_seg_or (when segmented); bit p of the posted mask is
// set when pass p reorders something (key passes first, then segment
// passes).  One workgroup.
constexpr int kOrPostThreads = 1024;
__global__ __launch_bounds__(kOrPostThreads) void k_or_post(const uint32_t *__restrict__ parts, int64_t nparts,
                                                           const uint32_t *__restrict__ seg_words, int key_passes,
                                                           int seg_passes, Mailbox *mb, unsigned long long seq) {
    __shared__ uint32_t s_or[kOrPostThreads / kWave][2];
    const int tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
    uint32_t v[2] = {0, 0};
    for (int64_t t = tid; t < nparts; t += kOrPostThreads) {
        const uint2 q = *(const uint2 *)(parts + 2 * t);
        v[0] |= q.x;
        v[1] |= q.y;
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const uint32_t r = wave_reduce(v[k], OrU());
        if (lane == 0) s_or[w][k] = r;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t o[2] = {0, 0};
        for (int i = 0; i < kOrPostThreads / kWave; i++)
            for (int k = 0; k < 2; k++) o[k] |= s_or[i][k];
        const uint64_t kdiff = ((uint64_t)o[1] << 32) | o[0];
        uint32_t so = 0, sc = 0;
        if (seg_words)
            for (int c = 0; c < kOrCopies; c++) {
                so |= seg_words[(c * kOrWords + 0) * kOrStride];
                sc |= seg_words[(c * kOrWords + 1) * kOrStride];
            }
        const uint32_t sdiff = so & sc;
        unsigned long long mask = 0;
        for (int p = 0; p < key_passes; p++)
            if ((kdiff >> (8 * p)) & 255u) mask |= 1ull << p;
        for (int q = 0; q < seg_passes; q++)
            if ((sdiff >> (8 * q)) & 255u) mask |= 1ull << (key_passes + q);
        mailbox_post(mb, seq, mask);
    }
}

__global__ __launch_bounds__(kBlock) void k_seg_or(const int *__restrict__ lens, int64_t m,
                                                   unsigned *__restrict__ words) {
    uint32_t o = 0, c = 0;
    for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < m; g += (int64_t)gridDim.x * kBlock)
        if (lens[g] != 0) {
            o |= (uint32_t)g;
            c |= ~(uint32_t)g;
        }
    const uint32_t v[2] = {o, c};
    or_publish<2>(v, words, 0);
}

// Segment-number digits: segment g contributes lens[g] elements.
__global__ void k_seg_hist(const int *__restrict__ lens, int64_t m, int npass, unsigned *__restrict__ hist) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < m; g += (int64_t)gridDim.x * blockDim.x) {
        unsigned L = (unsigned)lens[g];
        if (!L) continue;
        for (int p = 0; p < npass; p++) atomicAdd(hist + p * kRadix + ((g >> (8 * p)) & 255), L);
    }
}

// Exclusive scan of each pass's 256 counts (one block per pass).
__global__ void k_hist_scan(unsigned *hist, unsigned *starts) {
    __shared__ unsigned sh[kRadix];
    const int p = blockIdx.x, d = threadIdx.x;
    unsigned v = hist[p * kRadix + d];
    sh[d] = v;
    __syncthreads();
    for (int o = 1; o < kRadix; o <<= 1) {
        unsigned t = d >= o ? sh[d - o] : 0;
        __syncthreads();
        sh[d] += t;
        __syncthreads();
    }
    starts[p * kRadix + d] = sh[d] - v;
}

// ---------------------------------------------------------------- one pass
struct PassArgs {
    const void *src;         // first pass: the original values; else keys_in
    bool vec4;               // keys of 4 bytes or less (source 0, 2, 5, 6), read four at a time
    const int *idx_in;       // null on the first pass (index = position)
    void *keys_out;          // null when the next passes need no keys
    int kout;                // bytes per key in keys_out (8, 4, 2 or 1)
    int oshift;              // keys_out holds key >> oshift
    int *idx_out;            // the element order after this pass
    const unsigned *starts;  // 256 bucket starts of this pass
    const unsigned *tile_ofs;  // tiles x 256: elements of each digit in earlier tiles
    uint32_t *chain;           // or null: tiles x 256 chain words (chain_exclusive)
    unsigned *ticket;          // the chain's tile counter
    int parity;                // the chain's state set
    int64_t n;
    int shift;               // digit = (field >> shift) & 255
    // 0: int values, 1: double values, 2: keys32, 3: keys64, 4: segment
    // number, 5: keys16, 6: keys8.  Stored keys keep only the bits the
    // remaining passes read, shifted down: after the second of four 8-bit
    // passes a 32-bit key needs 16 bits, after the third 8.
    int source;
    int up;
    SegdView seg;                  // segment of an element (segmented ranks)
    unsigned *err;                 // device error word
    // The last pass of an unsegmented rank, fused with the first partition
    // pass of the inverse (partition_invert_begin): instead of the order it
    // writes the pairs (index, position) into child index >> clog of
    // `pairs`, at the tile's slot from tile_ofs2 (tiles x 256, the column
    // scan of the per-tile child counts that k_rank_hist wrote to hist2).
    uint2 *pairs;
    const unsigned *tile_ofs2;
    int clog;
};

// Key of element e for this pass (ix = its original index).
template <class KeyT>
__device__ __forceinline__ KeyT pass_key(const PassArgs &a, int64_t e, int ix) {
    if (a.source == 0) return (KeyT)int_key(((const int *)a.src)[e], a.up);
    if (a.source == 1) return (KeyT)dbl_key(((const double *)a.src)[e], a.up);
    if (a.source == 2) return (KeyT)((const uint32_t *)a.src)[e];
    if (a.source == 3) return (KeyT)((const uint64_t *)a.src)[e];
    if (a.source == 5) return (KeyT)((const uint16_t *)a.src)[e];
    if (a.source == 6) return (KeyT)((const uint8_t *)a.src)[e];
    return (KeyT)(uint32_t)a.seg.seg_of(ix);
}

// Four consecutive stored fields from e (a multiple of 4; the arrays are
// 256-byte aligned, the original values 16-byte aligned when vec4 is set),
// widened to 32 bits: the raw int for source 0, else the stored key.
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef unsigned char u8x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4 load_fields4(const PassArgs &a, int64_t e) {
    if (a.source == 5) {
        const u16x4 h = *(const u16x4 *)((const uint16_t *)a.src + e);
        return u32x4{h.x, h.y, h.z, h.w};
    }
    if (a.source == 6) {
        const u8x4 c = *(const u8x4 *)((const uint8_t *)a.src + e);
        return u32x4{c.x, c.y, c.z, c.w};
    }
    return *(const u32x4 *)((const uint32_t *)a.src + e);
}

// keys_out[pos] = key >> oshift, in KOUT bytes.
template <class KeyT, int KOUT>
__device__ __forceinline__ void store_key(const PassArgs &a, int64_t pos, KeyT k) {
    const KeyT v = k >> a.oshift;
    if constexpr (KOUT == 1) ((uint8_t *)a.keys_out)[pos] = (uint8_t)v;
    else if constexpr (KOUT == 2) ((uint16_t *)a.keys_out)[pos] = (uint16_t)v;
    else ((KeyT *)a.keys_out)[pos] = v;
}

// Per-tile digit counts of one pass, as a tiles x 256 matrix (tile t =
// elements [t*kRankTile, (t+1)*kRankTile) of the current order).
// With B2 (the fused last pass) it also counts each tile's elements per
// child of the inverse's first partition pass (index >> a.clog) into hist2.
// With ORW (the first pass, launched before the host knows which passes
// run) it also writes the bits in which the tile's keys differ from key 0
// to or_parts[2 tile], [2 tile + 1] (low, high 32 bits) for k_or_post.
template <class KeyT, bool B2, bool ORW = false>
__global__ __launch_bounds__(kRankThreads) void k_rank_hist(PassArgs a, unsigned *__restrict__ hist,
                                                            unsigned *__restrict__ hist2,
                                                            uint32_t *__restrict__ or_parts = nullptr) {
    __shared__ unsigned s_hist[kRadix];
    __shared__ unsigned s_hist2[B2 ? kRadix : 1];
    const int tid = threadIdx.x;
    // ORW: key 0 of the vector, the reference every tile compares with
    // (read first, so that its latency hides behind the tile's loads)
    KeyT k0 = 0;
    if constexpr (ORW) k0 = pass_key<KeyT>(a, 0, 0);
    for (int i = tid; i < kRadix; i += kRankThreads) {
        s_hist[i] = 0;
        if constexpr (B2) s_hist2[i] = 0;
    }
    __syncthreads();
    const int64_t tile = blockIdx.x;      // rows are whole 1 KiB stores: launch order
    const int64_t t0 = tile * kRankTile;
    // Item j is element t0 + ((j / 4) * kRankThreads + tid) * 4 + j % 4: runs
    // of four per lane, one 16-byte load each for 4-byte keys.
    static_assert(kRankItems % 4 == 0, "runs of four keys per lane");
    KeyT kv[kRankItems];
    bool in[kRankItems];
#pragma unroll
    for (int g = 0; g < kRankItems / 4; g++) {
        const int64_t e0 = t0 + ((int64_t)g * kRankThreads + tid) * 4;
        if (a.vec4 && e0 + 4 <= a.n) {
            const u32x4 w = load_fields4(a, e0);
            const uint32_t x[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int c = 0; c < 4; c++) {
                kv[g * 4 + c] = (KeyT)(a.source == 0 ? int_key((int)x[c], a.up) : x[c]);
                in[g * 4 + c] = true;
            }
        } else {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int64_t e = e0 + c;
                const int j = g * 4 + c;
                in[j] = e < a.n;
                kv[j] = 0;
                if (in[j]) kv[j] = pass_key<KeyT>(a, e, a.source == 4 ? (a.idx_in ? a.idx_in[e] : (int)e) : 0);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kRankItems; j++)
        if (in[j]) atomicAdd(&s_hist[(kv[j] >> a.shift) & 255], 1u);
    if constexpr (ORW) {
        // the bits in which some key of the tile differs from key 0 of the
        // vector (the same reference in every tile): a digit takes one value
        // over all keys iff none of its bits differs anywhere
        constexpr int NW = sizeof(KeyT) / 4;
        __shared__ uint32_t s_or[kRankThreads / kWave][NW];
        KeyT x = 0;
#pragma unroll
        for (int j = 0; j < kRankItems; j++)
            if (in[j]) x |= kv[j] ^ k0;
        const int w = tid / kWave, lane = tid % kWave;
#pragma unroll
        for (int k = 0; k < NW; k++) {
            const uint32_t r = wave_reduce((uint32_t)((uint64_t)x >> (32 * k)), OrU());
            if (lane == 0) s_or[w][k] = r;
        }
        __syncthreads();
        if (tid < 2) {
            uint32_t r = 0;
            if (tid < NW)
#pragma unroll
                for (int i = 0; i < kRankThreads / kWave; i++) r |= s_or[i][tid];
            or_parts[2 * tile + tid] = r;
        }
    }
    if constexpr (B2) {
        // the elements' original indices: children of the inverse
#pragma unroll
        for (int g = 0; g < kRankItems / 4; g++) {
            const int64_t e0 = t0 + ((int64_t)g * kRankThreads + tid) * 4;
            int ix[4];
            if (a.idx_in && e0 + 4 <= a.n) {
                const u32x4 w = *(const u32x4 *)((const uint32_t *)a.idx_in + e0);
                ix[0] = (int)w.x; ix[1] = (int)w.y; ix[2] = (int)w.z; ix[3] = (int)w.w;
            } else {
#pragma unroll
                for (int c = 0; c < 4; c++) ix[c] = e0 + c < a.n ? (a.idx_in ? a.idx_in[e0 + c] : (int)(e0 + c)) : -1;
            }
#pragma unroll
            for (int c = 0; c < 4; c++)
                if (e0 + c < a.n && (unsigned)ix[c] < (unsigned)a.n) atomicAdd(&s_hist2[(unsigned)ix[c] >> a.clog], 1u);
        }
    }
    __syncthreads();
    for (int d = tid; d < kRadix; d += kRankThreads) {
        hist[tile * kRadix + d] = s_hist[d];
        if constexpr (B2) hist2[tile * kRadix + d] = s_hist2[d];
    }
}

// KOUT: bytes per key written to keys_out (a.kout), 0 when no keys are.
// FUSE (KOUT 0): the pass writes the inverse's first partition pass
// (a.pairs) instead of the order.
template <class KeyT, int KOUT, bool FUSE = false>
__global__ __launch_bounds__(kRankThreads) void k_rank_pass(PassArgs a) {
    static_assert(!FUSE || KOUT == 0, "the fused last pass writes no keys");
    constexpr int W = kRankThreads / kWave;
    // s_cnt: per-wave digit counts, then per-wave starts within a digit;
    // s_stage: keys, then (reinterpreted) indices, in digit order.  FUSE:
    // the 8-byte pairs in child order, staged once every position is known,
    // over both (then dead): 32 KiB instead of 8 + 32, so three workgroups
    // of LDS become four.
    constexpr int kCntBytes = W * kRadix * 4;
    constexpr int kStageBytes = kRankTile * (int)sizeof(KeyT);
    constexpr int kPairBytes = kRankTile * 8;
    constexpr int kRaw = (FUSE && kPairBytes > kCntBytes + kStageBytes) ? kPairBytes : kCntBytes + kStageBytes;
    __shared__ __attribute__((aligned(16))) unsigned char s_raw[kRaw];
    unsigned(*const s_cnt)[kRadix] = (unsigned(*)[kRadix])s_raw;
    KeyT *const s_stage = (KeyT *)(s_raw + kCntBytes);
    __shared__ unsigned s_tot[kRadix];
    __shared__ unsigned s_lstart[kRadix];     // tile-local start of each digit's run
    __shared__ unsigned s_gbase[kRadix];      // global position of the run
    __shared__ unsigned char s_dig[FUSE ? 1 : kRankTile];
    __shared__ unsigned s_c2[FUSE ? kRadix : 1];   // FUSE: per-child counts, then starts, of the tile
    __shared__ unsigned s_g2[FUSE ? kRadix : 1];   // FUSE: the tile's slot in each child
    __shared__ unsigned s_placed;                  // FUSE: pairs staged
    __shared__ unsigned s_tile;

    const int tid = threadIdx.x;
    const int w = tid / kWave, lane = tid % kWave;
    for (int i = tid; i < W * kRadix; i += kRankThreads) (&s_cnt[0][0])[i] = 0;
    if constexpr (FUSE)
        for (int i = tid; i < kRadix; i += kRankThreads) s_c2[i] = 0;
    // a chained pass numbers its tiles in the order they start; the
    // order of the tiles' runs within a digit is then their data order
    if (kRankChain && a.chain && tid == 0) s_tile = atomicAdd(a.ticket, 1u);
    __syncthreads();
    const int64_t tile = (kRankChain && a.chain) ? (int64_t)s_tile : xcd_tile(blockIdx.x, gridDim.x);
    const int64_t t0 = tile * kRankTile;
    const int64_t count = a.n - t0 < kRankTile ? a.n - t0 : kRankTile;
    const int64_t wb = t0 + (int64_t)w * kWaveSpan;

    KeyT key[kRankItems];
    int idx[kRankItems];
    unsigned dig[kRankItems];
    unsigned loc[kRankItems];

    // Item j of a lane is element wb + j*64 + lane: each round of the
    // multisplit below holds 64 consecutive elements, which keeps it stable.
    // A whole span of 4-byte keys is read with 16-byte loads (lane l gets
    // elements 4l..4l+3 of each 256) and put in that order through the
    // wave's own part of s_stage, which is free until the staging phase.
    if constexpr (sizeof(KeyT) == 4) {
        if (a.vec4 && wb + kWaveSpan <= a.n) {
            static_assert(kRankItems % 4 == 0, "runs of four keys per lane");
            uint32_t *sw = (uint32_t *)s_stage + (int64_t)w * kWaveSpan;
            auto transpose = [&](auto load4, uint32_t (&out)[kRankItems]) {
#pragma unroll
                for (int g = 0; g < kRankItems / 4; g++) {
                    const u32x4 v = load4(wb + g * 4 * kWave + 4 * lane);
                    *(u32x4 *)(sw + g * 4 * kWave + 4 * lane) = v;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                for (int j = 0; j < kRankItems; j++) out[j] = sw[j * kWave + lane];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            };
            uint32_t raw[kRankItems];
            transpose([&](int64_t e) { return load_fields4(a, e); }, raw);
#pragma unroll
            for (int j = 0; j < kRankItems; j++) key[j] = (KeyT)(a.source == 0 ? int_key((int)raw[j], a.up) : raw[j]);
            if (a.idx_in) {
                transpose([&](int64_t e) { return *(const u32x4 *)((const uint32_t *)a.idx_in + e); }, raw);
#pragma unroll
                for (int j = 0; j < kRankItems; j++) idx[j] = (int)raw[j];
            } else {
#pragma unroll
                for (int j = 0; j < kRankItems; j++) idx[j] = (int)(wb + (int64_t)j * kWave + lane);
            }
            goto loaded;
        }
    }
#pragma unroll
    for (int j = 0; j < kRankItems; j++) {
        const int64_t e = wb + (int64_t)j * kWave + lane;
        KeyT k = 0;
        int ix = 0;
        if (e < a.n) {
            ix = a.idx_in ? a.idx_in[e] : (int)e;
            k = pass_key<KeyT>(a, e, ix);
        }
        key[j] = k;
        idx[j] = ix;
    }
loaded:
    // FULL: every element of the wave's span exists (all but the last
    // wave of the vector), so no lane is masked out.
    auto multisplit = [&](auto full) {
        constexpr bool FULL = decltype(full)::value;
#pragma unroll
        for (int j = 0; j < kRankItems; j++) {
            const bool in = FULL || wb + (int64_t)j * kWave + lane < a.n;
            const KeyT k = key[j];
            const unsigned d = in ? (unsigned)((k >> a.shift) & 255) : 0u;
            dig[j] = d;
            // lanes of this round holding the same digit: those that agree
            // with this lane in every bit of the digit, the AND over the
            // bits of (lanes with the bit set) XNOR (all lanes if this lane
            // has it set, none if not).  Per bit four instructions: the
            // sign-extended bit, its lane mask (a compare of that value,
            // which the compiler would otherwise rebuild from a shift of the
            // digit), and one 3-input bit operation (v_bitop3, gfx950) per
            // half of the mask: acc & ~(spread ^ mask), truth table 0x84
            // over (spread, acc, mask).
            uint32_t plo = ~0u, phi = ~0u;
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const uint32_t spread = (uint32_t)__builtin_amdgcn_sbfe((int)d, b, 1);   // 0 or ~0: bit b
                const uint64_t bb = __builtin_amdgcn_uicmp(spread, 0u, 33);            // lanes with bit b (ICMP_NE)
                plo = __builtin_amdgcn_bitop3_b32(spread, plo, (uint32_t)bb, 0x84);
                phi = __builtin_amdgcn_bitop3_b32(spread, phi, (uint32_t)(bb >> 32), 0x84);
            }
            if constexpr (!FULL) {             // lanes past the end are no one's peers
                const uint64_t act = ballot(in);
                plo &= (uint32_t)act;
                phi &= (uint32_t)(act >> 32);
            }
            // peers in lanes below this one (v_mbcnt: the set bits of a
            // 64-lane mask below the current lane)
            const unsigned before = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
            const uint64_t peers = ((uint64_t)phi << 32) | plo;
            const unsigned old = s_cnt[w][d];
            loc[j] = old + before;
            // the highest lane of each digit group advances the counter
            const bool leader = in && (peers >> lane) == 1ull;
            if (leader) s_cnt[w][d] = old + before + 1;
        }
    };
    if (wb + kWaveSpan <= a.n) multisplit(std::true_type{});
    else multisplit(std::false_type{});
    __syncthreads();

    // Per digit: wave starts within the digit, the tile count, and the
    // global start of the tile's run (bucket start + earlier tiles).
    if (tid < kRadix) {
        const int d = tid;
        unsigned run = 0;
#pragma unroll
        for (int i = 0; i < W; i++) {
            const unsigned c = s_cnt[i][d];
            s_cnt[i][d] = run;
            run += c;
        }
        s_tot[d] = run;
        const unsigned prefix = (kRankChain && a.chain)
                                    ? chain_exclusive(a.chain, tile, kRadix, d, run, a.parity, a.err)
                                    : a.tile_ofs[tile * kRadix + d];
        s_gbase[d] = a.starts[d] + prefix;
    }
    __syncthreads();
    // tile-local digit starts: exclusive scan of the 256 totals (one wave)
    if (tid < kWave) {
        unsigned c[kRadix / kWave], t = 0;
#pragma unroll
        for (int i = 0; i < kRadix / kWave; i++) { c[i] = s_tot[tid * (kRadix / kWave) + i]; t += c[i]; }
        unsigned run = wave_inclusive_scan(t, AddU()) - t;
#pragma unroll
        for (int i = 0; i < kRadix / kWave; i++) {
            const int d = tid * (kRadix / kWave) + i;
            s_lstart[d] = run;
            s_gbase[d] -= run;       // now: global position of local slot 0 of digit d's run
            run += c[i];
        }
    }
    __syncthreads();

    if constexpr (FUSE) {
        // Every element's sorted position is known here: place the pair
        // (index, position) in child index >> clog of the inverse's first
        // partition pass, staged in child order, each child's run written
        // contiguously at the tile's slot in that child.
        uint2 *const s_pair = (uint2 *)s_raw;   // written only after the barrier below: s_cnt is read here
        unsigned ch[kRankItems], rk[kRankItems], pos[kRankItems];
#pragma unroll
        for (int j = 0; j < kRankItems; j++) {
            const int64_t e = wb + (int64_t)j * kWave + lane;
            ch[j] = 0xFFFFFFFFu;
            if (e < a.n) {
                const unsigned d = dig[j];
                pos[j] = s_gbase[d] + s_lstart[d] + s_cnt[w][d] + loc[j];
                if ((unsigned)idx[j] < (unsigned)a.n) {
                    ch[j] = (unsigned)idx[j] >> a.clog;
                    rk[j] = atomicAdd(&s_c2[ch[j]], 1u);
                } else {
                    atomicOr(a.err, 1u);
                }
            }
        }
        __syncthreads();
        if (tid < kWave) {          // child starts within the tile (one wave)
            unsigned c[kRadix / kWave], t = 0;
#pragma unroll
            for (int i = 0; i < kRadix / kWave; i++) { c[i] = s_c2[tid * (kRadix / kWave) + i]; t += c[i]; }
            unsigned run = wave_inclusive_scan(t, AddU()) - t;
#pragma unroll
            for (int i = 0; i < kRadix / kWave; i++) {
                s_c2[tid * (kRadix / kWave) + i] = run;
                run += c[i];
            }
            if (tid == kWave - 1) s_placed = run;
        }
        for (int i = tid; i < kRadix; i += kRankThreads) s_g2[i] = a.tile_ofs2[tile * kRadix + i];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kRankItems; j++)
            if (ch[j] != 0xFFFFFFFFu) s_pair[s_c2[ch[j]] + rk[j]] = make_uint2((unsigned)idx[j], pos[j]);
        __syncthreads();
        // s_pair[0 .. placed): the tile's elements with a valid index, by
        // child (index < n < 2^(clog + 8), so every child is below 256)
        const int placed = (int)s_placed;
        const int64_t csize = (int64_t)1 << a.clog;
        for (int i = tid; i < placed; i += kRankThreads) {
            const uint2 pr = s_pair[i];
            const unsigned c = pr.x >> a.clog;
            const int64_t at = (int64_t)c * csize + s_g2[c] + (i - s_c2[c]);
            const int64_t cend = ((int64_t)c + 1) * csize < a.n ? ((int64_t)c + 1) * csize : a.n;
            if (at < cend) a.pairs[at] = pr;
            else atomicOr(a.err, 1u);
        }
        return;
    }

    // Stage in digit order, then write each digit's run contiguously:
    // first the keys (when a later pass needs them), then the indices.
    unsigned slot[kRankItems];
#pragma unroll
    for (int j = 0; j < kRankItems; j++) {
        const int64_t e = wb + (int64_t)j * kWave + lane;
        slot[j] = 0;
        if (e < a.n) {
            slot[j] = s_lstart[dig[j]] + s_cnt[w][dig[j]] + loc[j];
            s_dig[slot[j]] = (unsigned char)dig[j];
            if constexpr (KOUT > 0) s_stage[slot[j]] = key[j];
        }
    }
    __syncthreads();
    // Staged element i = tid + k * kRankThreads goes to s_gbase[its digit]
    // + i, for the keys and then for the indices.  (s_gbase[d] >= 0: the
    // tile's elements below digit d are among all elements below it.)
    unsigned spos[kRankItems];
#pragma unroll
    for (int k = 0; k < kRankItems; k++) {
        const int i = tid + k * kRankThreads;
        spos[k] = i < count ? s_gbase[s_dig[i]] + (unsigned)i : 0xFFFFFFFFu;
    }
    if constexpr (KOUT > 0) {
#pragma unroll
        for (int k = 0; k < kRankItems; k++) {
            const int i = tid + k * kRankThreads;
            if (i < count) {
                if (spos[k] < a.n) store_key<KeyT, KOUT>(a, spos[k], s_stage[i]);
                else atomicOr(a.err, 1u);
            }
        }
        __syncthreads();
    }
    int *s_idx = (int *)s_stage;
#pragma unroll
    for (int j = 0; j < kRankItems; j++) {
        const int64_t e = wb + (int64_t)j * kWave + lane;
        if (e < a.n) s_idx[slot[j]] = idx[j];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kRankItems; k++) {
        const int i = tid + k * kRankThreads;
        if (i < count) {
            if (spos[k] < a.n) a.idx_out[spos[k]] = s_idx[i];
            else atomicOr(a.err, 1u);
        }
    }
}

// rank[k] = k - (segment offset): the result when no pass reorders anything.
__global__ void k_rank_identity(int *rank, int64_t n, SegdView seg, int segmented) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        rank[k] = (int)(segmented ? k - seg.offs[seg.seg_of(k)] : k);
}

// ---------------------------------------------------------------- host driver
struct RankScratch {
    int64_t keysA, keysB, idxA, idxB, hist, starts, counters, status, scan_status, status2, orparts, invert, bytes;
    RankScratch(int64_t n, bool dbl) {
        auto up = [](int64_t b) { return (b + 255) & ~(int64_t)255; };
        int64_t kw = dbl ? 8 : 4;
        int64_t tiles = (n + kRankTile - 1) / kRankTile;
        keysA = 0;
        keysB = up(keysA + kw * n);
        idxA = up(keysB + kw * n);
        idxB = up(idxA + 4 * n);
        hist = up(idxB + 4 * n);
        starts = up(hist + 4 * 12 * kRadix);
        counters = up(starts + 4 * 12 * kRadix);
        status = up(counters + 4 * 16);
        scan_status = up(status + 4 * kRadix * (tiles > 0 ? tiles : 1));
        // one column-scan status region per pass (at most 12), one for the
        // fused last pass's child counts and one for the speculative first
        // pass, cleared together
        status2 = up(scan_status + 14 * 8 * tile_scan_status_words(tiles > 0 ? tiles : 1));
        // the fused last pass's per-tile child counts; the first pass's
        // per-tile OR words
        orparts = up(status2 + 4 * kRadix * (tiles > 0 ? tiles : 1));
        invert = up(orparts + 8 * (tiles > 0 ? tiles : 1));
        bytes = up(invert + partition_permute_scratch_bytes(n, 4));
    }
};

static unsigned *g_hist_host = nullptr;

void rank_run(int *rank, const void *s, int64_t n, bool dbl, bool up, const void *sd, int64_t m, void *scratch) {
    if (n <= 0) return;
    if (!scratch) usage_error("rank called without scratch");
    if (n > 0x7fffffffLL) usage_error("rank: too many keys");
    RankScratch L(n, dbl);
    char *b = (char *)scratch;
    unsigned *hist = (unsigned *)(b + L.hist);
    unsigned *starts = (unsigned *)(b + L.starts);
    unsigned *tile_ofs = (unsigned *)(b + L.status);
    uint64_t *scan_status = (uint64_t *)(b + L.scan_status);
    const bool segmented = sd != nullptr && m > 1;
    SegdView seg(sd, n, segmented ? m : 0);
    const int key_passes = dbl ? 8 : 4;
    int seg_passes = 0;
    if (segmented) {
        int64_t mm = m - 1;
        while (mm > 0) { seg_passes++; mm >>= 8; }
    }
    const int total = key_passes + seg_passes;
    const int64_t tiles = (n + kRankTile - 1) / kRankTile;

    const bool chain = kRankChain && n < (int64_t)kChainMask;
    const int64_t sw = tile_scan_status_words(tiles);
    if (!g_hist_host) HIPCHECK(hipHostMalloc(&g_hist_host, 4 * 12 * kRadix, hipHostMallocDefault));
    const bool aligned = ((uintptr_t)s & 15) == 0;
    int g = grid_for(n, kBlock * 16);
    if (g > 2048) g = 2048;
    int gs = grid_for(m, kBlock);
    if (gs > 4096) gs = 4096;
    bool runs[12] = {};
    // The speculative first histogram (below): pass 0's tile counts, column
    // scanned with its own status region (index 13).
    unsigned *tile_ofs2 = (unsigned *)(b + L.status2);
    InvertFirstPass inv{};
    bool spec = false;
    if (chain) {
        // chained passes need every digit's start up front: full histograms
        fill_u32(hist, 0, 12 * kRadix);
        if (dbl) k_key_hist<true><<<g, kBlock, 0, g_stream>>>(s, n, up, aligned, hist);
        else k_key_hist<false><<<g, kBlock, 0, g_stream>>>(s, n, up, aligned, hist);
        LAUNCHCHECK();
        if (seg_passes) {
            k_seg_hist<<<gs, kBlock, 0, g_stream>>>(seg.lens, m, seg_passes, hist + key_passes * kRadix);
            LAUNCHCHECK();
        }
        k_hist_scan<<<total, kRadix, 0, g_stream>>>(hist, starts);
        LAUNCHCHECK();
        HIPCHECK(hipMemcpyAsync(g_hist_host, hist, 4 * (size_t)total * kRadix, hipMemcpyDeviceToHost, g_stream));
        HIPCHECK(hipStreamSynchronize(g_stream));
        for (int p = 0; p < total; p++) {
            int nonzero = 0;
            for (int d = 0; d < kRadix; d++) nonzero += g_hist_host[p * kRadix + d] != 0;
            runs[p] = nonzero > 1;
        }
    } else {
        // Pass 0's tile counts, which nearly always serve, also give the
        // keys' OR words; k_or_post posts which passes run to the host.
        // Behind it the device clears the scan status regions and the
        // inverse's counters and column-scans pass 0's counts, covering the
        // host's wait and its next launches.
        unsigned *words = hist;
        static_assert(kOrCopies * kOrWords * kOrStride <= 12 * kRadix, "the OR words fit the histogram space");
        if (seg_passes) {
            fill_u32(words, 0, kOrCopies * kOrWords * kOrStride);
            k_seg_or<<<gs, kBlock, 0, g_stream>>>(seg.lens, m, words);
            LAUNCHCHECK();
        }
        PassArgs a0{};
        a0.n = n;
        a0.up = up;
        a0.seg = seg;
        a0.source = dbl ? 1 : 0;
        a0.src = s;
        a0.shift = 0;
        a0.vec4 = !dbl && aligned;
        a0.err = device_error_word();
        uint32_t *or_parts = (uint32_t *)(b + L.orparts);
        if (dbl)
            k_rank_hist<uint64_t, false, true><<<(unsigned)tiles, kRankThreads, 0, g_stream>>>(a0, tile_ofs, nullptr,
                                                                                              or_parts);
        else
            k_rank_hist<uint32_t, false, true><<<(unsigned)tiles, kRankThreads, 0, g_stream>>>(a0, tile_ofs, nullptr,
                                                                                              or_parts);
        LAUNCHCHECK();
        const unsigned long long seq = mailbox_next();
        k_or_post<<<1, kOrPostThreads, 0, g_stream>>>(or_parts, tiles, seg_passes ? words : nullptr, key_passes,
                                                      seg_passes, mailbox_device(), seq);
        LAUNCHCHECK();
        fill_u32(scan_status, 0, 2 * sw * 14);      // status regions: 12 passes, the inverse's counts, pass 0's
        tile_scan(tile_ofs, tiles, nullptr, scan_status + 13 * sw, true, starts);
        spec = true;
        // An unsegmented rank whose inverse takes two or more partition
        // passes runs the first of them inside its last pass (FUSE): the pass
        // knows every element's sorted position as it places it, so the
        // order is never written and read back.
        if (!segmented) inv = partition_invert_begin(n, b + L.invert);
        const unsigned long long mask = mailbox_wait(seq);
        for (int p = 0; p < total; p++) runs[p] = (mask >> p) & 1u;
    }
    int run[12], nrun = 0;
    for (int p = 0; p < total; p++)
        if (runs[p]) run[nrun++] = p;
    if (nrun == 0) {
        k_rank_identity<<<grid_for(n, kBlock * 4), kBlock, 0, g_stream>>>(rank, n, seg, segmented);
        LAUNCHCHECK();
        return;
    }
    // The last key pass that runs decides whether keys must still travel.
    int last_key_run = -1;
    for (int i = 0; i < nrun; i++) if (run[i] < key_passes) last_key_run = i;
    // pass 0's speculative counts serve when pass 0 runs, unless it is also
    // the fused last pass, which counts the inverse's children as well
    spec = spec && run[0] == 0 && !(inv.pairs && nrun == 1);

    // Chained passes: per-digit counts below 2^30 (n < 2^30); one set of
    // words for all passes (alternating state sets) and a tile counter per
    // pass, contiguous and cleared together.
    unsigned *tickets = (unsigned *)(b + L.counters);
    if (chain) fill_u32(tickets, 0, (L.status - L.counters) / 4 + kRadix * tiles);
    void *keys[2] = {b + L.keysA, b + L.keysB};
    int *idxs[2] = {(int *)(b + L.idxA), (int *)(b + L.idxB)};
    int cur = -1;   // buffer holding the current (keys, idx); -1 = original input
    int kshift = 0;  // the stored keys are the transformed keys >> kshift
    int kbytes = dbl ? 8 : 4;
    for (int i = 0; i < nrun; i++) {
        const int p = run[i];
        const int out = cur == 0 ? 1 : 0;
        PassArgs a;
        a.n = n;
        a.up = up;
        a.seg = seg;
        a.starts = starts + p * kRadix;
        a.tile_ofs = tile_ofs;
        a.chain = chain ? (uint32_t *)tile_ofs : nullptr;
        a.ticket = tickets + i;
        a.parity = i & 1;
        a.idx_in = cur < 0 ? nullptr : idxs[cur];
        if (p < key_passes) {
            a.source = cur < 0 ? (dbl ? 1 : 0) : (dbl ? 3 : (kbytes == 4 ? 2 : (kbytes == 2 ? 5 : 6)));
            a.src = cur < 0 ? s : keys[cur];
            a.shift = 8 * p - kshift;
        } else {
            a.source = 4;
            a.src = nullptr;
            a.shift = 8 * (p - key_passes);
        }
        a.vec4 = a.source == 5 || a.source == 6 || ((a.source == 0 || a.source == 2) && ((uintptr_t)a.src & 15) == 0);
        a.keys_out = i < last_key_run ? keys[out] : nullptr;
        a.kout = kbytes;
        a.oshift = 0;
        if (a.keys_out && !dbl) {
            // the next pass that reads keys is run[i + 1] (key passes come first)
            const int nshift = 8 * run[i + 1];
            const int width = 32 - nshift;
            a.kout = width > 16 ? 4 : (width > 8 ? 2 : 1);
            a.oshift = nshift - kshift;
        }
        a.idx_out = idxs[out];
        a.err = device_error_word();
        const bool fuse = inv.pairs && i + 1 == nrun;
        a.pairs = fuse ? inv.pairs : nullptr;
        a.tile_ofs2 = tile_ofs2;
        a.clog = inv.clog;
        if (!chain && !(i == 0 && spec)) {
            if (fuse) {
                if (dbl) k_rank_hist<uint64_t, true><<<(unsigned)tiles, kRankThreads, 0, g_stream>>>(a, tile_ofs, tile_ofs2);
                else k_rank_hist<uint32_t, true><<<(unsigned)tiles, kRankThreads, 0, g_stream>>>(a, tile_ofs, tile_ofs2);
            } else {
                if (dbl) k_rank_hist<uint64_t, false><<<(unsigned)tiles, kRankThreads, 0, g_stream>>>(a, tile_ofs, nullptr);
                else k_rank_hist<uint32_t, false><<<(unsigned)tiles, kRankThreads, 0, g_stream>>>(a, tile_ofs, nullptr);
            }
            LAUNCHCHECK();
            // the column scan also leaves the digits' starts (the scan of its totals)
            tile_scan(tile_ofs, tiles, nullptr, scan_status + i * sw, true, starts + p * kRadix);
            // the children's totals go to the inverse's counters
            if (fuse) tile_scan(tile_ofs2, tiles, inv.cnt, scan_status + 12 * sw, true);
        }
        const dim3 grid((unsigned)tiles), block(kRankThreads);
        if (fuse) {
            if (dbl) k_rank_pass<uint64_t, 0, true><<<grid, block, 0, g_stream>>>(a);
            else k_rank_pass<uint32_t, 0, true><<<grid, block, 0, g_stream>>>(a);
        } else if (dbl) {
            if (a.keys_out) k_rank_pass<uint64_t, 8><<<grid, block, 0, g_stream>>>(a);
            else k_rank_pass<uint64_t, 0><<<grid, block, 0, g_stream>>>(a);
        } else if (!a.keys_out) {
            k_rank_pass<uint32_t, 0><<<grid, block, 0, g_stream>>>(a);
        } else if (a.kout == 1) {
            k_rank_pass<uint32_t, 1><<<grid, block, 0, g_stream>>>(a);
        } else if (a.kout == 2) {
            k_rank_pass<uint32_t, 2><<<grid, block, 0, g_stream>>>(a);
        } else {
            k_rank_pass<uint32_t, 4><<<grid, block, 0, g_stream>>>(a);
        }
        LAUNCHCHECK();
        if (a.keys_out) {
            kshift += a.oshift;
            kbytes = a.kout;
        }
        cur = out;
    }
    // rank = inverse of the final order
    if (inv.pairs) partition_invert_finish(rank, n, b + L.invert);
    else partition_invert(rank, idxs[cur], n, segmented ? sd : nullptr, m, b + L.invert);
}

int64_t rank_scratch_bytes(int64_t n, bool dbl) { return RankScratch(n, dbl).bytes; }

}  // namespace cvlhip

using namespace cvlhip;

#define RANKS(UD, UPV, TY, DBL)                                                                  \
    void rk##UD##_lu##TY(vec_p d, vec_p s, int len, vec_p scratch) {                             \
        rank_run((int *)d, s, len, DBL, UPV, nullptr, 0, scratch);                               \
    }                                                                                            \
    int rk##UD##_lu##TY##_scratch(int len) { return units_for_bytes(rank_scratch_bytes(len, DBL)); } \
    unsigned int rk##UD##_lu##TY##_inplace(void) { return INPLACE_NONE; }                        \
    void rk##UD##_le##TY(vec_p d, vec_p s, vec_p segd, int vec_len, int seg_count, vec_p scratch) { \
        rank_run((int *)d, s, vec_len, DBL, UPV, segd, seg_count, scratch);                      \
    }                                                                                            \
    int rk##UD##_le##TY##_scratch(int vec_len, int seg_count) {                                  \
        (void)seg_count;                                                                         \
        return units_for_bytes(rank_scratch_bytes(vec_len, DBL));                                \
    }                                                                                            \
    unsigned int rk##UD##_le##TY##_inplace(void) { return INPLACE_NONE; }

extern "C" {
RANKS(u, true, z, false)
RANKS(d, false, z, false)
RANKS(u, true, d, true)
RANKS(d, false, d, true)
}
