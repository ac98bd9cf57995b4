// scan_kernel.hpp -- the single-pass look-back scan kernel, shared by the
// CVL scans and reductions (scan.hip) and by generated fused operations
// whose result is a segmented reduction of an expression (the input then
// comes from an expression loader instead of an array).
#pragma once

#include "cvl_hip.hpp"
#include "lookback.hpp"
#include "segd.hpp"

namespace cvlhip {

template <class T> struct Vec16;
template <> struct Vec16<int> { typedef int type __attribute__((ext_vector_type(4))); };
template <> struct Vec16<double> { typedef double type __attribute__((ext_vector_type(2))); };

// Elements each lane holds per row: one 16-byte load's worth, by default.
template <class Op> struct lane_values { static constexpr int value = 16 / (int)sizeof(typename Op::T); };

// VB: bytes each lane holds per row; 16 (one 16-byte load) unless the
// caller asks for wider lanes (the flat segmented-reduction tiles).
template <class Op, int R, int B = kBlock, int VB = 16>
struct ScanShape {
    using T = typename Op::T;
    static_assert(VB % 16 == 0, "whole 16-byte loads per lane");
    static constexpr int V = VB == 16 ? lane_values<Op>::value : VB / (int)sizeof(T);   // elements per lane per row
    static constexpr int ROW = kWave * V;            // elements per wave row
    static constexpr int WAVES = B / kWave;
    static constexpr int WAVE_ELEMS = R * ROW;
    static constexpr int TILE = WAVES * WAVE_ELEMS;
};

template <class Op> struct ordered { static constexpr bool value = false; };
template <> struct ordered<MaxD> { static constexpr bool value = true; };
template <> struct ordered<MinD> { static constexpr bool value = true; };

#ifndef CVLHIP_SCAN_ROWS
#define CVLHIP_SCAN_ROWS 16
#endif
constexpr int kScanRows = CVLHIP_SCAN_ROWS;   // rows per wave in one tile

// Load V elements starting at element `base` (neutral beyond n): whole
// 16-byte loads when aligned and V is a multiple of a load's elements.
template <class Op, int V>
__device__ __forceinline__ void load_lane(const typename Op::T *__restrict__ in, int64_t base, int64_t n,
                                          bool aligned, typename Op::T (&x)[V]) {
    using T = typename Op::T;
    constexpr int E = 16 / (int)sizeof(T);          // elements per 16-byte load
    constexpr bool VEC = V % E == 0;                // (callers with fewer elements pass aligned = false)
    if (VEC && aligned && base + V <= n) {
        typename Vec16<T>::type v[VEC ? V / E : 1];
#pragma unroll
        for (int q = 0; q < V / E; q++) v[q] = *(const typename Vec16<T>::type *)(in + base + q * E);
#pragma unroll
        for (int q = 0; q < V / E; q++)
#pragma unroll
            for (int k = 0; k < E; k++) x[q * E + k] = load_val<Op>(v[q][k]);
    } else {
#pragma unroll
        for (int k = 0; k < V; k++) x[k] = base + k < n ? load_val<Op>(in[base + k]) : Op::neutral();
    }
}

template <class T, int V>
__device__ __forceinline__ void store_lane(T *__restrict__ out, int64_t base, int64_t n, bool aligned,
                                           const T (&y)[V]) {
    constexpr int E = 16 / (int)sizeof(T);          // elements per 16-byte store
    static_assert(V % E == 0, "whole 16-byte stores per lane");
    if (aligned && base + V <= n) {
#pragma unroll
        for (int q = 0; q < V / E; q++) {
            typename Vec16<T>::type v;
#pragma unroll
            for (int k = 0; k < E; k++) v[k] = y[q * E + k];
            *(typename Vec16<T>::type *)(out + base + q * E) = v;
        }
    } else {
#pragma unroll
        for (int k = 0; k < V; k++)
            if (base + k < n) out[base + k] = y[k];
    }
}

// Inclusive wave scan of (flag, value) pairs under the segmented operator
// (fa, a) . (fb, b) = (fa | fb, fb ? b : a op b).
template <class Op, class T>
__device__ __forceinline__ void seg_wave_scan(int &f, T &v, Op op) {
    const T e = Op::neutral();
#define CVLHIP_SEGSTEP(CTRL, RM)                          \
    {                                                     \
        T pv = dpp_move<CTRL, RM>(e, v);                  \
        int pf = dpp_move<CTRL, RM>(0, f);                \
        v = f ? v : op(pv, v);                            \
        f = f | pf;                                       \
    }
    CVLHIP_SEGSTEP(kRowShr1, 0xf)
    CVLHIP_SEGSTEP(kRowShr2, 0xf)
    CVLHIP_SEGSTEP(kRowShr4, 0xf)
    CVLHIP_SEGSTEP(kRowShr8, 0xf)
    CVLHIP_SEGSTEP(kRowBcast15, 0xa)
    CVLHIP_SEGSTEP(kRowBcast31, 0xc)
#undef CVLHIP_SEGSTEP
}

struct SegBits {
    const uint32_t *heads;   // null when unsegmented
    const int *wcnt;
    const int *nz;
    const void *sd;          // when set (and heads null): read the rest from this descriptor's header
    int64_t m;               // segment count, when known (0: unknown); lets reductions see that nz is the identity
};

// Where a scan's input comes from: an array (ArrayIn), or -- in generated
// fused operations -- an expression evaluated per element.
template <class T>
struct ArrayIn {
    const T *p;
    template <class Op, int V>
    __device__ __forceinline__ void load(int64_t base, int64_t n, bool aligned, T (&x)[V]) const {
        load_lane<Op, V>(p, base, n, aligned, x);
    }
};

// Per-tile records of a segmented reduction run without look-back (NOLB):
// each tile reduces with an empty incoming prefix and records
//   agg[t], flag[t]: its aggregate under the segmented operator (flag: the
//                    tile holds a segment head; agg is then the value from
//                    its last head to its end);
//   slot[t], val[t]: the output slot of the segment that enters the tile
//                    from earlier tiles and ends in it (-1 if none), and
//                    that segment's partial value within the tile.
// k_seg_reduce_fix then combines each such partial with the exclusive
// prefix of the tile.  Laid out from `base` for g tiles.
template <class T>
struct NolbRecords {
    T *agg;
    int *flag;
    int *slot;
    T *val;
    // each array starts a multiple of 16 bytes after `base`
    __host__ __device__ static int64_t up16(int64_t b) { return (b + 15) & ~(int64_t)15; }
    __host__ __device__ static int64_t flag_at(int64_t g) { return up16((int64_t)sizeof(T) * g); }
    __host__ __device__ static int64_t slot_at(int64_t g) { return up16(flag_at(g) + 4 * g); }
    __host__ __device__ static int64_t val_at(int64_t g) { return up16(slot_at(g) + 4 * g); }
    __host__ __device__ static int64_t bytes(int64_t g) { return val_at(g) + (int64_t)sizeof(T) * g; }
    __host__ __device__ static NolbRecords at(void *base, int64_t g) {
        char *b = (char *)base;
        NolbRecords r;
        r.agg = (T *)b;
        r.flag = (int *)(b + flag_at(g));
        r.slot = (int *)(b + slot_at(g));
        r.val = (T *)(b + val_at(g));
        return r;
    }
};

// One kernel for: unsegmented exclusive scan (SEG=false, REDUCE=false),
// segmented exclusive scan (SEG=true, REDUCE=false) and segmented reduce
// (SEG=true, REDUCE=true; `out` has one element per segment).
// NOLB (segmented reductions only): no look-back; `status` holds the
// NolbRecords of gridDim.x tiles and k_seg_reduce_fix finishes the
// segments that cross tiles.
//
// scan_tile is the work of one workgroup on one tile (`tile` of `ntiles`;
// NOLB records are laid out for ntiles tiles).  k_scan runs one tile per
// workgroup; the segmented-op schedules (k_seg_sched) loop over tiles.

template <class Op, int R, bool SEG, bool REDUCE, class In, int B, bool NOLB, int VB = 16>
__device__ __forceinline__ void scan_tile(const In &in, typename Op::T *__restrict__ out, int64_t n, SegBits sb,
                                          uint64_t *__restrict__ status, int64_t tile, int64_t ntiles,
                                          bool aligned, unsigned *__restrict__ err) {
    using T = typename Op::T;
    using S = ScanShape<Op, R, B, VB>;
    constexpr int V = S::V;
    static_assert(V <= 16, "end flags in bits 0..15 of eb, the word's earlier heads above");
    constexpr uint32_t VMASK = (1u << V) - 1u;
    static_assert(!NOLB || (SEG && REDUCE), "only segmented reductions run without look-back");
    static_assert(VB == 16 || (SEG && REDUCE), "wide lanes only in segmented reductions");
    const Op op;

    __shared__ T s_wagg[S::WAVES];
    __shared__ int s_wflag[S::WAVES];
    __shared__ T s_win[S::WAVES];
    __shared__ int s_winf[S::WAVES];   // NOLB: a head in an earlier wave of the tile
    __shared__ int s_fix_slot;         // NOLB: this tile's record of the segment entering it
    __shared__ T s_fix_val;

    if (NOLB && threadIdx.x == 0) {    // read by thread 0 only after the barriers below
        s_fix_slot = -1;
        s_fix_val = Op::neutral();
    }
    if constexpr (SEG) {
        if (!sb.heads) {
            const SegdView v = SegdView::from_header(sb.sd);
            sb.heads = v.heads;
            sb.wcnt = v.wcnt;
            sb.nz = v.nz;
            sb.m = v.m;
        }
    }
    const int w = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int64_t wbase = tile * S::TILE + (int64_t)w * S::WAVE_ELEMS + (int64_t)lane * V;

    T x[R][V];
    uint32_t hb[R];
    uint32_t eb[R];
    // Reductions: the heads at or before element k of a lane's row part are
    // wc[r] (heads before the part's bitmap word, loaded with the data and
    // first used at the stores) + the word's heads before the part (kept in
    // bits 16.. of eb[r]) + the part's own heads up to k.  With no empty
    // segments the output slot of a segment is that count - 1 and nz need
    // not be read.
    constexpr bool WC = REDUCE;
    int wc[WC ? R : 1];
    bool nz_ident = false;
    if constexpr (WC) {
        if (sb.m > 0) nz_ident = sb.wcnt[(n + 31) >> 5] == sb.m;
    }
#pragma unroll
    for (int r = 0; r < R; r++) in.template load<Op, V>(wbase + (int64_t)r * S::ROW, n, aligned, x[r]);
    // The tile's head bitmap words (and, for reductions, head counts) are
    // staged in LDS with one coalesced pass, so the rows below read them
    // there instead of with two more vector loads each.  Words past the
    // descriptor's last (index (n+31)/32) read as 0.
    constexpr int HW = (int)(S::TILE / 32) + 2;
    __shared__ uint32_t s_heads[SEG ? HW : 1];
    __shared__ int s_wcnt[WC ? HW : 1];
    const int64_t w0 = (tile * S::TILE) >> 5;
    if constexpr (SEG) {
        const int64_t W = (n + 31) >> 5;
        __syncthreads();                    // a caller looping over tiles is done with the last one's
        for (int i = threadIdx.x; i < HW; i += B) {
            const int64_t wd = w0 + i;
            s_heads[i] = wd <= W ? sb.heads[wd] : 0u;
            if constexpr (WC) s_wcnt[i] = wd <= W ? sb.wcnt[wd] : 0;
        }
        __syncthreads();
    }
    auto head_word = [&](int64_t wd) -> uint32_t { return s_heads[wd - w0]; };
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t base = wbase + (int64_t)r * S::ROW;
        hb[r] = 0;
        eb[r] = 0;
        if constexpr (WC) wc[r] = 0;
        if constexpr (SEG) {
            if (base < n) {
                const int64_t wd = base >> 5;
                const int sh = (int)(base & 31);
                const uint64_t two = (uint64_t)head_word(wd) | ((uint64_t)head_word(wd + 1) << 32);
                hb[r] = (uint32_t)(two >> sh) & VMASK;
                if constexpr (REDUCE) {
                    uint32_t nxt = (uint32_t)(two >> (sh + 1)) & VMASK;   // head at element + 1
                    if (base + V >= n) {                                  // the last element ends its segment
                        int64_t last = n - 1 - base;
                        if (last < V) nxt |= 1u << last;
                        nxt &= (last + 1 >= 32) ? 0xffffffffu : ((1u << (last + 1)) - 1u);
                    }
                    eb[r] = nxt;
                    if constexpr (WC) {
                        eb[r] |= (uint32_t)__popc((uint32_t)two & ((1u << sh) - 1u)) << 16;
                        if (nxt) wc[r] = s_wcnt[wd - w0];
                    }
                }
            }
        }
    }

    // Wave-local scan of the wave's R rows.
    T carry = Op::neutral();
    int cf = 0;
    uint64_t needs = 0;     // bit r*V+k: element still needs the incoming prefix
    static_assert(R * V <= 64, "needs mask holds one bit per element of a lane");
#pragma unroll
    for (int r = 0; r < R; r++) {
        T a = Op::neutral();
        int af = 0;
#pragma unroll
        for (int k = 0; k < V; k++) {
            if (SEG && ((hb[r] >> k) & 1u)) { a = x[r][k]; af = 1; }
            else a = op(a, x[r][k]);
        }
        // A row of the wave with no segment head (most rows of long
        // segments) scans without the flags: the same steps and bits.
        if constexpr (SEG) {
            if (ballot(hb[r] != 0u)) seg_wave_scan(af, a, op);
            else a = wave_inclusive_scan(a, op);
        } else {
            a = wave_inclusive_scan(a, op);
        }
        const T ra = bcast_last(a);
        const int rf = SEG ? bcast_last(af) : 0;
        // A reduction needs per-element values only where a segment ends:
        // a row of the wave with no end only moves the carry on.
        bool elems = true;
        if constexpr (REDUCE) elems = ballot((eb[r] & VMASK) != 0u) != 0;
        if (elems) {
            T ea = shfl_up1(a);
            int ef = SEG ? shfl_up1(af) : 0;
            if (lane == 0) { ea = Op::neutral(); ef = 0; }
            T p = ef ? ea : op(carry, ea);
            int pf = cf | ef;
#pragma unroll
            for (int k = 0; k < V; k++) {
                const bool h = SEG && ((hb[r] >> k) & 1u);
                T y;
                if constexpr (!REDUCE) {
                    if (h) { y = Op::neutral(); p = x[r][k]; pf = 1; }
                    else { y = p; if (!pf) needs |= 1ull << (r * V + k); p = op(p, x[r][k]); }
                } else {
                    if (h) { p = x[r][k]; pf = 1; }
                    else { p = op(p, x[r][k]); if (!pf) needs |= 1ull << (r * V + k); }
                    y = p;
                }
                x[r][k] = y;
            }
        }
        carry = rf ? ra : op(carry, ra);
        cf |= rf;
    }

    if (lane == 0) { s_wagg[w] = carry; s_wflag[w] = cf; }
    __syncthreads();

    if (w == 0) {
        T tv = Op::neutral();
        int tf = 0;
        T wpv[S::WAVES];
        int wpf[S::WAVES];
#pragma unroll
        for (int i = 0; i < S::WAVES; i++) {
            wpv[i] = tv;
            wpf[i] = tf;
            const T a = s_wagg[i];
            const int f = s_wflag[i];
            tv = f ? a : op(tv, a);
            tf |= f;
        }
        T P = Op::neutral();
        if constexpr (NOLB) {
            if (lane == 0) {
                const NolbRecords<T> rec = NolbRecords<T>::at(status, ntiles);
                rec.agg[tile] = tv;
                rec.flag[tile] = tf;
            }
        } else if (tile == 0) {
            using St = TileStatus<T>;
            if (lane == 0) St::publish(status, 0, kPrefix, tv);
        } else {
            using St = TileStatus<T>;
            const bool done = SEG && tf;      // a head inside the tile: later tiles need nothing earlier
            if (lane == 0) St::publish(status, tile, done ? kPrefix : kAggregate, tv);
            bool first_head = false;
            if constexpr (SEG) {
                const int64_t t0 = tile * S::TILE;
                first_head = (head_word(t0 >> 5) >> (t0 & 31)) & 1u;
            }
            if (!first_head) P = lookback(status, tile, op, err);
            if (!done && lane == 0) St::publish(status, tile, kPrefix, op(P, tv));
        }
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < S::WAVES; i++) {
                s_win[i] = wpf[i] ? wpv[i] : op(P, wpv[i]);
                s_winf[i] = wpf[i];
            }
        }
    }
    __syncthreads();

    const T inc = s_win[w];
    const T cid = Op::cvl_id();
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t base = wbase + (int64_t)r * S::ROW;
        T y[V], pv[NOLB ? V : 1];
#pragma unroll
        for (int k = 0; k < V; k++) {
            T v = x[r][k];
            if ((needs >> (r * V + k)) & 1ull) v = op(inc, v);
            if constexpr (NOLB) pv[k] = v;
            y[k] = op(cid, v);
        }
        if constexpr (!REDUCE) {
            store_lane<T, V>(out, base, n, aligned, y);
        } else {
#pragma unroll
            for (int k = 0; k < V; k++) {
                if ((eb[r] >> k) & 1u) {
                    const int rk = wc[r] + (int)(eb[r] >> 16) + __popc(hb[r] & ((2u << k) - 1u));
                    const int slot = nz_ident ? rk - 1 : sb.nz[rk - 1];
                    if (NOLB && ((needs >> (r * V + k)) & 1ull) && !s_winf[w]) {
                        // the end of the segment that entered the tile: its
                        // value still lacks the tile's incoming prefix
                        s_fix_slot = slot;
                        if constexpr (NOLB) s_fix_val = pv[k];
                    } else {
                        out[slot] = y[k];
                    }
                }
            }
        }
    }
    if constexpr (NOLB) {
        __syncthreads();
        if (threadIdx.x == 0) {
            const NolbRecords<T> rec = NolbRecords<T>::at(status, ntiles);
            rec.slot[tile] = s_fix_slot;
            rec.val[tile] = s_fix_val;
        }
    }
}

template <class Op, int R, bool SEG, bool REDUCE, class In = ArrayIn<typename Op::T>, int B = kBlock,
          bool NOLB = false, int VB = 16>
__global__ __launch_bounds__(B) void k_scan(In in, typename Op::T *__restrict__ out, int64_t n, SegBits sb,
                                            uint64_t *__restrict__ status, unsigned *__restrict__ counter,
                                            int aligned_i, unsigned *__restrict__ err) {
    __shared__ unsigned s_tile;
    // tile ids in launch order (look-back waits only on started tiles); a
    // single-tile launch passes no counter
    if (threadIdx.x == 0) s_tile = counter ? atomicAdd(counter, 1u) : blockIdx.x;
    __syncthreads();
    scan_tile<Op, R, SEG, REDUCE, In, B, NOLB, VB>(in, out, n, sb, status, (int64_t)s_tile, (int64_t)gridDim.x,
                                                   aligned_i != 0, err);
}

// Rows per wave, threads per workgroup and bytes per lane and row of the
// flat 8-row-arm tiles of a segmented reduction (the CVL entries and
// generated fused reductions, so that both give the same bits); the
// autotuner picks them (2 rows x 256 threads x 64 bytes: 4096 doubles or
// 8192 ints per tile).
#ifndef CVLHIP_SEG_REDUCE_ROWS
#define CVLHIP_SEG_REDUCE_ROWS 2
#endif
constexpr int kSegReduceRows = CVLHIP_SEG_REDUCE_ROWS;
// Bytes each lane holds per row in those tiles: 64 (8 doubles, 16 ints) by
// default.  A lane folds its consecutive elements in order before the wave
// scan of the row, so the tile runs a quarter of the wave scans of 16-byte
// lanes for the same elements (the scans were most of its instructions).
#ifndef CVLHIP_SEG_REDUCE_LANE_BYTES
#define CVLHIP_SEG_REDUCE_LANE_BYTES 64
#endif
constexpr int kSegReduceLaneBytes = CVLHIP_SEG_REDUCE_LANE_BYTES;
// Rows per wave of the largest tiles of segmented scans (the flat8 arm).
constexpr int kSegScanRows = 8;
#ifndef CVLHIP_SEG_REDUCE_BLOCK
#define CVLHIP_SEG_REDUCE_BLOCK 512
#endif
constexpr int kSegReduceBlock = CVLHIP_SEG_REDUCE_BLOCK;

// Second step of a NOLB segmented reduction (one workgroup of kFixThreads):
// the exclusive prefix of every tile under the segmented operator, folded
// in tile order, completes the segment that entered the tile:
// out[slot[t]] = cvl_id op (prefix(t) op val[t]).
constexpr int kFixThreads = 1024;
constexpr int kFixPer = 8;   // consecutive records per thread and round

// Records j0 .. j0+kFixPer-1 of one array (`fill` past g): 16-byte loads
// when all of them exist and the array is 16-byte aligned (j0 is a
// multiple of kFixPer).
template <class T>
__device__ __forceinline__ void fix_load(const T *__restrict__ p, int64_t j0, int64_t g, bool vec, T fill,
                                         T (&x)[kFixPer]) {
    typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
    typedef unsigned long long u64x2_t __attribute__((ext_vector_type(2)));
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "four- or eight-byte records");
    if (vec && j0 + kFixPer <= g) {
        if constexpr (sizeof(T) == 4) {
#pragma unroll
            for (int q = 0; q < kFixPer / 4; q++) {
                const u32x4_t w = *(const u32x4_t *)(p + j0 + 4 * q);
                const unsigned e[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const unsigned u = e[c];
                    x[4 * q + c] = __builtin_bit_cast(T, u);
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < kFixPer / 2; q++) {
                const u64x2_t w = *(const u64x2_t *)(p + j0 + 2 * q);
                const unsigned long long e[2] = {w.x, w.y};
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    const unsigned long long u = e[c];
                    x[2 * q + c] = __builtin_bit_cast(T, u);
                }
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < kFixPer; k++) x[k] = j0 + k < g ? p[j0 + k] : fill;
    }
}

template <class Op>
__device__ __forceinline__ void seg_reduce_fix_body(NolbRecords<typename Op::T> rec, int64_t g,
                                                    typename Op::T *__restrict__ out) {
    using T = typename Op::T;
    constexpr int W = kFixThreads / kWave;
    constexpr int64_t ROUND = (int64_t)kFixThreads * kFixPer;
    const Op op;
    __shared__ T s_v[W];
    __shared__ int s_f[W];
    const int tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
    const T cid = Op::cvl_id();
    const bool vec = (((uintptr_t)rec.agg | (uintptr_t)rec.flag | (uintptr_t)rec.slot | (uintptr_t)rec.val) & 15) == 0;
    // Rounds of ROUND records (one round up to 8192 tiles).  Each thread
    // folds its kFixPer consecutive records in registers; one segmented scan
    // across the workgroup, entered with the fold of all earlier rounds
    // (carry, the same in every thread), gives each thread what precedes its
    // first record; the thread then walks its records again and completes
    // the segments that entered them.  That is a few instructions per
    // record, against a wave scan (hundreds of instructions) per 64 records
    // when each thread holds one: the one workgroup is bound by issue.
    // (16 records per thread were slower: each 16-byte load of a wave then
    // spans 8 KiB, twice the cache lines.)
    T carry = Op::neutral();
    int cf = 0;
    for (int64_t base = 0; base < g; base += ROUND) {
        const int64_t j0 = base + (int64_t)tid * kFixPer;
        T a[kFixPer], pv[kFixPer];
        int f[kFixPer], sl[kFixPer];
        fix_load(rec.agg, j0, g, vec, Op::neutral(), a);
        fix_load(rec.flag, j0, g, vec, 0, f);
        fix_load(rec.slot, j0, g, vec, -1, sl);
        fix_load(rec.val, j0, g, vec, Op::neutral(), pv);
        T iv = Op::neutral();
        int iff = 0;
#pragma unroll
        for (int k = 0; k < kFixPer; k++) {
            iv = f[k] ? a[k] : op(iv, a[k]);
            iff |= f[k];
        }
        seg_wave_scan(iff, iv, op);
        if (lane == kWave - 1) { s_v[w] = iv; s_f[w] = iff; }
        __syncthreads();
        // the fold through each wave: wp stops before this wave's first
        // thread, tp runs through the round's last
        T tp = carry, wp = carry;
        int tf = cf;
#pragma unroll
        for (int i = 0; i < W; i++) {
            if (i == w) wp = tp;
            tp = s_f[i] ? s_v[i] : op(tp, s_v[i]);
            tf |= s_f[i];
        }
        T ev = shfl_up1(iv);
        int ef = shfl_up1(iff);
        if (lane == 0) { ev = Op::neutral(); ef = 0; }
        T p = ef ? ev : op(wp, ev);                 // everything before this thread's first record
#pragma unroll
        for (int k = 0; k < kFixPer; k++) {
            if (sl[k] >= 0) out[sl[k]] = op(cid, op(p, pv[k]));
            p = f[k] ? a[k] : op(p, a[k]);
        }
        carry = tp;
        cf = tf;
        __syncthreads();                            // s_v is written again next round
    }
}

template <class Op>
__global__ __launch_bounds__(kFixThreads) void k_seg_reduce_fix(NolbRecords<typename Op::T> rec, int64_t g,
                                                                typename Op::T *__restrict__ out) {
    seg_reduce_fix_body<Op>(rec, g, out);
}


// Writes the identity for every empty segment (segmented reductions).
template <class T>
__global__ void k_empty_identity(T *__restrict__ out, const int *__restrict__ lens, int64_t m, T id) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; g < m; g += (int64_t)gridDim.x * blockDim.x)
        if (lens[g] == 0) out[g] = id;
}


// The same for a descriptor known only on the device (its header gives m).
template <class T>
__global__ void k_empty_identity_hdr(T *__restrict__ out, const void *__restrict__ sd, T id) {
    const SegdView v = SegdView::from_header(sd);
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; g < v.m; g += (int64_t)gridDim.x * blockDim.x)
        if (v.lens[g] == 0) out[g] = id;
}

}  // namespace cvlhip
