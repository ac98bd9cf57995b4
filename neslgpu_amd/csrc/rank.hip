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
// Each reordering pass is one "onesweep" kernel: a tile of 4096 elements
// ranks its digits stably with a wave-level multisplit (8 ballots give each
// lane the lanes holding the same digit), gets each digit's count in earlier
// tiles from a per-digit decoupled look-back, and scatters (key, index)
// straight to the next buffer; the last pass writes rank[index] instead.
#include "cvl_hip.hpp"
#include "segd.hpp"
#include "lookback.hpp"
#include "../../include/cvl.h"

namespace cvlhip {

constexpr int kRadix = 256;
constexpr int kRankItems = 16;                           // per lane
constexpr int kRankTile = kBlock * kRankItems;           // 4096
constexpr int kWaveSpan = kWave * kRankItems;            // 1024 elements per wave

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
template <bool DBL>
__global__ __launch_bounds__(kBlock) void k_key_hist(const void *__restrict__ s, int64_t n, bool up,
                                                     unsigned *__restrict__ hist) {
    constexpr int P = DBL ? 8 : 4;
    __shared__ unsigned sh[P][kRadix];
    for (int i = threadIdx.x; i < P * kRadix; i += kBlock) (&sh[0][0])[i] = 0;
    __syncthreads();
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock) {
        uint64_t key = DBL ? dbl_key(((const double *)s)[k], up) : int_key(((const int *)s)[k], up);
#pragma unroll
        for (int p = 0; p < P; p++) atomicAdd(&sh[p][(key >> (8 * p)) & 255], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < P * kRadix; i += kBlock)
        if ((&sh[0][0])[i]) atomicAdd(hist + i, (&sh[0][0])[i]);
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
    const int *idx_in;       // null on the first pass (index = position)
    void *keys_out;          // null when the next passes need no keys
    int *idx_out;            // null on the last pass
    int *rank;               // non-null on the last pass
    const unsigned *starts;  // 256 bucket starts of this pass
    uint64_t *status;        // tiles x 256 look-back words
    unsigned *counter;
    int64_t n;
    int shift;               // digit = (field >> shift) & 255
    int source;              // 0: int values, 1: double values, 2: keys32, 3: keys64, 4: segment number
    int up;
    uint32_t flag_agg, flag_pre;   // look-back flags of this pass
    SegdView seg;                  // segment of an element (segmented ranks)
    int subtract_offset;           // last pass of a segmented rank
    unsigned *err;                 // device error word
};

template <class KeyT>
__global__ __launch_bounds__(kBlock) void k_rank_pass(PassArgs a) {
    __shared__ unsigned s_tile;
    __shared__ unsigned s_cnt[kBlock / kWave][kRadix];
    __shared__ unsigned s_base[kBlock / kWave][kRadix];   // per wave, per digit: global start

    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (threadIdx.x == 0) s_tile = atomicAdd(a.counter, 1u);
    for (int i = threadIdx.x; i < (kBlock / kWave) * kRadix; i += kBlock) (&s_cnt[0][0])[i] = 0;
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t wb = tile * kRankTile + (int64_t)w * kWaveSpan;

    KeyT key[kRankItems];
    int idx[kRankItems];
    unsigned dig[kRankItems];
    unsigned loc[kRankItems];
    const uint64_t lt = (lane == 0) ? 0 : (~0ull >> (64 - lane));

#pragma unroll
    for (int j = 0; j < kRankItems; j++) {
        const int64_t e = wb + (int64_t)j * kWave + lane;
        const bool in = e < a.n;
        KeyT k = 0;
        int ix = 0;
        if (in) {
            ix = a.idx_in ? a.idx_in[e] : (int)e;
            if (a.source == 0) k = (KeyT)int_key(((const int *)a.src)[e], a.up);
            else if (a.source == 1) k = (KeyT)dbl_key(((const double *)a.src)[e], a.up);
            else if (a.source == 2) k = (KeyT)((const uint32_t *)a.src)[e];
            else if (a.source == 3) k = (KeyT)((const uint64_t *)a.src)[e];
            else k = (KeyT)(uint32_t)a.seg.seg_of(ix);
        }
        key[j] = k;
        idx[j] = ix;
        const unsigned d = in ? (unsigned)((k >> a.shift) & 255) : 0u;
        dig[j] = d;
        // lanes of this round holding the same digit (inactive lanes excluded)
        uint64_t peers = ballot(in);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t bb = ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const unsigned before = (unsigned)__popcll(peers & lt);
        const unsigned old = s_cnt[w][d];
        loc[j] = old + before;
        // the highest lane of each digit group advances the counter
        const bool leader = in && (peers >> lane) == 1ull;
        if (leader) s_cnt[w][d] = old + before + 1;
    }
    __syncthreads();

    // Per digit: counts of the 4 waves, the tile count, and the look-back.
    {
        const int d = threadIdx.x;     // kBlock == kRadix
        unsigned c[kBlock / kWave], tot = 0;
#pragma unroll
        for (int i = 0; i < kBlock / kWave; i++) { c[i] = tot; tot += s_cnt[i][d]; }
        uint64_t *st = a.status + tile * kRadix + d;
        unsigned prefix = 0;
        if (tile == 0) {
            st_status(st, ((uint64_t)a.flag_pre << 32) | tot);
        } else {
            st_status(st, ((uint64_t)a.flag_agg << 32) | tot);
            for (int64_t j = tile - 1; j >= 0; j--) {
                const uint64_t *q = a.status + j * kRadix + d;
                uint64_t v = ld_status(q);
                for (int spins = 0; (uint32_t)(v >> 32) < a.flag_agg; spins++) {
                    if (spins >= kSpinLimit) {           // corrupted status: give up, flag it
                        atomicOr(a.err, 2u);
                        v = ((uint64_t)a.flag_pre << 32);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    v = ld_status(q);
                }
                prefix += (uint32_t)v;
                if ((uint32_t)(v >> 32) == a.flag_pre) break;
            }
            st_status(st, ((uint64_t)a.flag_pre << 32) | (prefix + tot));
        }
        const unsigned g0 = a.starts[d] + prefix;
#pragma unroll
        for (int i = 0; i < kBlock / kWave; i++) s_base[i][d] = g0 + c[i];
    }
    __syncthreads();

#pragma unroll
    for (int j = 0; j < kRankItems; j++) {
        const int64_t e = wb + (int64_t)j * kWave + lane;
        if (e >= a.n) continue;
        const int64_t pos = (int64_t)s_base[w][dig[j]] + loc[j];
        if (pos < 0 || pos >= a.n) { atomicOr(a.err, 1u); continue; }
        if (a.rank) {
            int64_t r = pos;
            if (a.subtract_offset) r -= a.seg.offs[a.seg.seg_of(idx[j])];
            a.rank[idx[j]] = (int)r;
        } else {
            if (a.keys_out) ((KeyT *)a.keys_out)[pos] = key[j];
            a.idx_out[pos] = idx[j];
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
    int64_t keysA, keysB, idxA, idxB, hist, starts, counters, status, bytes;
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
        bytes = up(status + 8 * kRadix * (tiles > 0 ? tiles : 1));
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
    unsigned *counters = (unsigned *)(b + L.counters);
    uint64_t *status = (uint64_t *)(b + L.status);
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

    fill_u32(hist, 0, 12 * kRadix);
    fill_u32(counters, 0, 16);
    fill_u32(status, 0, 2 * kRadix * tiles);
    {
        int g = grid_for(n, kBlock * 16);
        if (g > 2048) g = 2048;
        if (dbl) k_key_hist<true><<<g, kBlock, 0, g_stream>>>(s, n, up, hist);
        else k_key_hist<false><<<g, kBlock, 0, g_stream>>>(s, n, up, hist);
        LAUNCHCHECK();
        if (seg_passes) {
            int gs = grid_for(m, kBlock);
            if (gs > 4096) gs = 4096;
            k_seg_hist<<<gs, kBlock, 0, g_stream>>>(seg.lens, m, seg_passes, hist + key_passes * kRadix);
            LAUNCHCHECK();
        }
        k_hist_scan<<<total, kRadix, 0, g_stream>>>(hist, starts);
        LAUNCHCHECK();
    }
    // Which passes reorder anything?  (A pass whose 256 counts hold one
    // non-zero bucket keeps the order and is skipped.)
    if (!g_hist_host) HIPCHECK(hipHostMalloc(&g_hist_host, 4 * 12 * kRadix, hipHostMallocDefault));
    HIPCHECK(hipMemcpyAsync(g_hist_host, hist, 4 * (size_t)total * kRadix, hipMemcpyDeviceToHost, g_stream));
    HIPCHECK(hipStreamSynchronize(g_stream));
    int run[12], nrun = 0;
    for (int p = 0; p < total; p++) {
        int nonzero = 0;
        for (int d = 0; d < kRadix; d++) nonzero += g_hist_host[p * kRadix + d] != 0;
        if (nonzero > 1) run[nrun++] = p;
    }
    if (nrun == 0) {
        k_rank_identity<<<grid_for(n, kBlock * 4), kBlock, 0, g_stream>>>(rank, n, seg, segmented);
        LAUNCHCHECK();
        return;
    }
    // The last key pass that runs decides whether keys must still travel.
    int last_key_run = -1;
    for (int i = 0; i < nrun; i++) if (run[i] < key_passes) last_key_run = i;

    void *keys[2] = {b + L.keysA, b + L.keysB};
    int *idxs[2] = {(int *)(b + L.idxA), (int *)(b + L.idxB)};
    int cur = -1;   // buffer holding the current (keys, idx); -1 = original input
    for (int i = 0; i < nrun; i++) {
        const int p = run[i];
        const bool last = i == nrun - 1;
        const int out = cur == 0 ? 1 : 0;
        PassArgs a;
        a.n = n;
        a.up = up;
        a.seg = seg;
        a.starts = starts + p * kRadix;
        a.status = status;
        a.counter = counters + i;
        a.flag_agg = 2 * i + 1;
        a.flag_pre = 2 * i + 2;
        a.idx_in = cur < 0 ? nullptr : idxs[cur];
        if (p < key_passes) {
            a.source = cur < 0 ? (dbl ? 1 : 0) : (dbl ? 3 : 2);
            a.src = cur < 0 ? s : keys[cur];
            a.shift = 8 * p;
        } else {
            a.source = 4;
            a.src = nullptr;
            a.shift = 8 * (p - key_passes);
        }
        a.keys_out = (!last && i < last_key_run) ? keys[out] : nullptr;
        a.idx_out = last ? nullptr : idxs[out];
        a.rank = last ? rank : nullptr;
        a.subtract_offset = last && segmented;
        a.err = device_error_word();
        if (dbl) k_rank_pass<uint64_t><<<(unsigned)tiles, kBlock, 0, g_stream>>>(a);
        else k_rank_pass<uint32_t><<<(unsigned)tiles, kBlock, 0, g_stream>>>(a);
        LAUNCHCHECK();
        cur = out;
    }
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
