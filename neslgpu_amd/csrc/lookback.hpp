// lookback.hpp -- tile status words and the wave-wide decoupled look-back
// used by the single-pass scan, segmented scan/reduce and pack kernels.
//
// Each tile publishes its aggregate (flag kAggregate) as soon as it knows it,
// and its inclusive prefix (flag kPrefix) once that is known.  A 32-bit value
// and its flag share one 64-bit word; a 64-bit value is split over two words
// that each carry the flag, and a reader accepts the pair only when both
// flags agree.  Flags only move forward (invalid -> aggregate -> prefix), so
// a reader that sees disagreeing flags simply reads again.
#pragma once

#include "cvl_hip.hpp"

namespace cvlhip {

enum : uint32_t { kInvalid = 0, kAggregate = 1, kPrefix = 2 };

template <class T, int WORDS = (int)(sizeof(T) / 4)>
struct TileStatus;

template <class T>
struct TileStatus<T, 1> {
    static constexpr int kWords = 1;
    __device__ static void publish(uint64_t *st, int64_t t, uint32_t flag, T v) {
        st_status(st + t, ((uint64_t)flag << 32) | __builtin_bit_cast(uint32_t, v));
    }
    __device__ static uint32_t read(const uint64_t *st, int64_t t, T &v) {
        uint64_t w = ld_status(st + t);
        v = __builtin_bit_cast(T, (uint32_t)w);
        return (uint32_t)(w >> 32);
    }
};

// Values of two or more words: every word carries the flag; a reader
// accepts the value only when all flags agree.
template <class T, int WORDS>
struct TileStatus {
    static_assert(WORDS >= 2 && sizeof(T) == 4 * WORDS, "status of whole 32-bit words");
    static constexpr int kWords = WORDS;
    __device__ static void publish(uint64_t *st, int64_t t, uint32_t flag, T v) {
        const Words<T> u = __builtin_bit_cast(Words<T>, v);
        const uint64_t f = (uint64_t)flag << 32;
#pragma unroll
        for (int i = 0; i < WORDS; i++) st_status(st + WORDS * t + i, f | (uint32_t)u.w[i]);
    }
    __device__ static uint32_t read(const uint64_t *st, int64_t t, T &v) {
        Words<T> u;
        uint32_t f0 = 0;
        bool agree = true;
#pragma unroll
        for (int i = 0; i < WORDS; i++) {
            const uint64_t w = ld_status(st + WORDS * t + i);
            const uint32_t f = (uint32_t)(w >> 32);
            if (i == 0) f0 = f;
            agree = agree && f == f0;
            u.w[i] = (int)(uint32_t)w;
        }
        if (!agree) return kInvalid;
        v = __builtin_bit_cast(T, u);
        return f0;
    }
};

// Polls before a look-back gives up.  A valid launch never gets near it (every
// tile publishes before it looks back, and tiles start in order); it only
// guarantees that a corrupted status array cannot hang the GPU.  Giving up
// raises bit 1 of the device error word.
constexpr int kSpinLimit = 1 << 22;


// Exclusive prefix of tile `tile`, computed by one whole wave.  Each round
// inspects the 64*K tiles before `pos`: lane l reads tiles
// pos - l*K - j (j = 0..K-1, all loads in flight together).  The nearest
// tile holding a full prefix ends the walk; otherwise the window slides back.
// Values are combined in tile order (earliest operand first), so
// order-sensitive operators (double max/min with signed zeros) agree with a
// left fold.  Tiles before tile 0 read as a prefix equal to the neutral.
// Operators whose result depends on association (double + and *).  Their
// look-back folds the window strictly in tile order and never slides the
// window (it re-polls until a prefix appears in it), so every inclusive
// prefix equals the sequential fold of the tile aggregates: the result does
// not depend on timing and is the same on every run.
template <class Op> struct seq_combine { static constexpr bool value = false; };
template <> struct seq_combine<AddD> { static constexpr bool value = true; };
template <> struct seq_combine<MulD> { static constexpr bool value = true; };

// Tiles inspected by each lane per look-back round (the window is 64*K).
// A sliding look-back does best with one; a fixed window (seq_combine) must
// hold a tile with a full prefix before it can finish, and the prefix
// frontier trails the newest tiles by more than 64 of them at full
// occupancy, so it inspects two per lane.
#ifndef CVLHIP_LOOKBACK_K
#define CVLHIP_LOOKBACK_K 1
#endif
template <class Op>
struct lookback_k { static constexpr int value = seq_combine<Op>::value ? 2 : CVLHIP_LOOKBACK_K; };

template <class Op>
__device__ typename Op::T lookback(const uint64_t *__restrict__ st, int64_t tile, Op op,
                                   unsigned *__restrict__ err) {
    using T = typename Op::T;
    using St = TileStatus<T>;
    constexpr int K = lookback_k<Op>::value;
    const int lane = lane_id();
    T acc = Op::neutral();
    int64_t pos = tile - 1;
    for (int rounds = 0;; rounds++) {
        T v[K];
        uint32_t f[K];
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int64_t t = pos - (int64_t)lane * K - j;
            v[j] = Op::neutral();
            f[j] = t >= 0 ? kInvalid : kPrefix;
        }
        // Poll until every tile nearer than the nearest full prefix in the
        // window has published (tiles beyond that prefix do not matter, so
        // a straggler there is not waited for); with no prefix in the
        // window, until all of it has.
        for (int spins = 0;; spins++) {
            int pj = K, bj = K;          // nearest prefix / nearest unpublished tile of this lane
#pragma unroll
            for (int j = K - 1; j >= 0; j--) {
                const int64_t t = pos - (int64_t)lane * K - j;
                if (f[j] == kInvalid) f[j] = St::read(st, t, v[j]);
                if (f[j] == kPrefix) pj = j;
                if (f[j] == kInvalid) bj = j;
            }
            const uint64_t pmask = ballot(pj < K);
            const int PL = pmask ? (int)__builtin_ctzll(pmask) : kWave;
            const bool blocks = lane < PL ? bj < K : (lane == PL && bj < pj);
            if (ballot(blocks) == 0) break;
            if (spins >= kSpinLimit) {
#pragma unroll
                for (int j = 0; j < K; j++)
                    if (f[j] == kInvalid) { f[j] = kPrefix; v[j] = Op::neutral(); }
                if (lane == 0) atomicOr(err, 2u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        // nearest prefix within this lane's K tiles
        int fj = K;
#pragma unroll
        for (int j = K - 1; j >= 0; j--)
            if (f[j] == kPrefix) fj = j;
        const uint64_t pm = ballot(fj < K);
        const int L = pm ? (int)__builtin_ctzll(pm) : kWave;
        // lane-local fold, earliest tile first, over the tiles this lane contributes
        const int lim = lane < L ? K - 1 : (lane == L ? fj : -1);
        T c = Op::neutral();
#pragma unroll
        for (int j = K - 1; j >= 0; j--)
            if (j <= lim) c = op(c, v[j]);
        if constexpr (seq_combine<Op>::value) {
            if (!pm && rounds < kSpinLimit) {      // re-poll the same window
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            if (!pm && lane == 0) atomicOr(err, 2u);
            // Strict left fold from the prefix tile: the later tiles of lane
            // Ls, then those of lanes Ls-1 .. 0, each lane's earliest first.
            const int Ls = L < kWave ? L : kWave - 1;
            const int jL = L < kWave ? shfl(fj, Ls) : K - 1;
            T w = Op::neutral();
            bool first = true;
            for (int l = Ls; l >= 0; l--) {
#pragma unroll
                for (int j = K - 1; j >= 0; j--) {
                    if (l == Ls && j > jL) continue;
                    const T x = op(Op::neutral(), shfl(v[j], l));
                    w = first ? x : op(w, x);
                    first = false;
                }
            }
            return w;
        }
        T r = shfl(c, kWave - 1 - lane);          // lane 0 now holds the earliest tiles
        acc = op(wave_reduce(r, op), acc);
        if (pm) break;
        pos -= (int64_t)kWave * K;
    }
    return acc;
}

// ---------------------------------------------------------------- column chains
// Per-tile counts of many columns (digits, buckets) chained from tile to
// tile: one 32-bit word per (tile, column) holds a 2-bit state in bits
// 30-31 and a count in bits 0-29 (counts below 2^30), so a word is written
// and read whole.  Tile t publishes its own count (aggregate), adds up the
// words of tiles t-1, t-2, ... until one holds an inclusive prefix, and
// publishes its own inclusive prefix.  Tiles must be numbered in the order
// they started (claimed from a counter), so every tile waited on is running
// or done.
//
// Two state sets alternate between launches over the same words: every word
// ends a launch in the `pre` state of its set, which is the `invalid` state
// of the other set, so a buffer cleared to zero once serves any number of
// launches with parity 0, 1, 0, ...
struct ChainStates {
    uint32_t invalid, agg, pre;
};
__device__ __forceinline__ ChainStates chain_states(int parity) {
    return parity ? ChainStates{2u << 30, 3u << 30, 0u} : ChainStates{0u, 1u << 30, 2u << 30};
}
constexpr uint32_t kChainMask = (1u << 30) - 1;
constexpr int kChainDepth = 8;          // words in flight per step of a walk

// Called by the one thread of column d in tile t: publishes c and returns
// the column's sum over tiles 0..t-1.  `cols` words per tile.
__device__ __forceinline__ uint32_t chain_exclusive(uint32_t *words, int64_t t, int cols, int d, uint32_t c,
                                                    int parity, unsigned *err) {
    const ChainStates S = chain_states(parity);
    auto ld = [](const uint32_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); };
    auto st = [](uint32_t *p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); };
    if (t == 0) {
        st(words + d, S.pre | c);
        return 0;
    }
    st(words + t * cols + d, S.agg | c);
    uint32_t excl = 0;
    int64_t u = t - 1;
    int spins = 0;
    for (;;) {
        uint32_t w[kChainDepth];
#pragma unroll
        for (int j = 0; j < kChainDepth; j++) w[j] = u - j >= 0 ? ld(words + (u - j) * cols + d) : S.pre;
        int used = 0;
        bool done = false;
#pragma unroll
        for (int j = 0; j < kChainDepth; j++) {
            if (done || used < j) continue;          // stop at the first unpublished word
            const uint32_t f = w[j] & ~kChainMask;
            if (f == S.invalid) continue;
            excl += w[j] & kChainMask;
            used = j + 1;
            done = f == S.pre;
        }
        if (done) break;
        u -= used;
        if (used) {
            spins = 0;
        } else {
            if (++spins >= kSpinLimit) {
                atomicOr(err, 2u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    st(words + t * cols + d, S.pre | ((excl + c) & kChainMask));
    return excl;
}

}  // namespace cvlhip
