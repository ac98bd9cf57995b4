// scan.hip -- scans and reductions, unsegmented and segmented
// (ref nesl/cvl/serial/vprims.c:17-201).
//
// Scans are exclusive; segmented scans restart from the identity at every
// segment.  All of them are single-pass decoupled look-back scans: a tile of
// TILE elements is loaded once with 16-byte lane loads, scanned in registers
// (lane-sequential, wave-wide DPP scan per row, rows in order, waves
// combined through LDS), its exclusive prefix comes from the look-back over
// earlier tiles, and it is written once: 2 bytes moved per byte of input.
// Segmented scans carry a (head-seen, value) pair; heads come from the
// descriptor's bitmap, so no per-element segment ids are read.
// Segmented reductions run the same pass and write the inclusive value at
// each segment's last element.
#include "cvl_hip.hpp"
#include "lookback.hpp"
#include "segd.hpp"
#include "scan.hpp"
#include "../../include/cvl.h"

namespace cvlhip {

template <class T> struct Vec16;
template <> struct Vec16<int> { typedef int type __attribute__((ext_vector_type(4))); };
template <> struct Vec16<double> { typedef double type __attribute__((ext_vector_type(2))); };

template <class Op, int R>
struct ScanShape {
    using T = typename Op::T;
    static constexpr int V = 16 / (int)sizeof(T);   // elements per lane per row
    static constexpr int ROW = kWave * V;            // elements per wave row
    static constexpr int WAVES = kBlock / kWave;
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

// Load V elements starting at element `base` (neutral beyond n).
template <class Op, int V>
__device__ __forceinline__ void load_lane(const typename Op::T *__restrict__ in, int64_t base, int64_t n,
                                          bool aligned, typename Op::T (&x)[V]) {
    using T = typename Op::T;
    if (aligned && base + V <= n) {
        typename Vec16<T>::type v = *(const typename Vec16<T>::type *)(in + base);
#pragma unroll
        for (int k = 0; k < V; k++) x[k] = load_val<Op>(v[k]);
    } else {
#pragma unroll
        for (int k = 0; k < V; k++) x[k] = base + k < n ? load_val<Op>(in[base + k]) : Op::neutral();
    }
}

template <class T, int V>
__device__ __forceinline__ void store_lane(T *__restrict__ out, int64_t base, int64_t n, bool aligned,
                                           const T (&y)[V]) {
    if (aligned && base + V <= n) {
        typename Vec16<T>::type v;
#pragma unroll
        for (int k = 0; k < V; k++) v[k] = y[k];
        *(typename Vec16<T>::type *)(out + base) = v;
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
};

// One kernel for: unsegmented exclusive scan (SEG=false, REDUCE=false),
// segmented exclusive scan (SEG=true, REDUCE=false) and segmented reduce
// (SEG=true, REDUCE=true; `out` has one element per segment).
template <class Op, int R, bool SEG, bool REDUCE>
__global__ __launch_bounds__(kBlock) void k_scan(const typename Op::T *__restrict__ in,
                                                 typename Op::T *__restrict__ out, int64_t n, SegBits sb,
                                                 uint64_t *__restrict__ status, unsigned *__restrict__ counter,
                                                 int aligned_i, unsigned *__restrict__ err) {
    using T = typename Op::T;
    using S = ScanShape<Op, R>;
    using St = TileStatus<T>;
    constexpr int V = S::V;
    constexpr uint32_t VMASK = (1u << V) - 1u;
    const Op op;
    const bool aligned = aligned_i != 0;

    __shared__ unsigned s_tile;
    __shared__ T s_wagg[S::WAVES];
    __shared__ int s_wflag[S::WAVES];
    __shared__ T s_win[S::WAVES];

    // tile ids in launch order (look-back waits only on started tiles); a
    // single-tile launch passes no counter
    if (threadIdx.x == 0) s_tile = counter ? atomicAdd(counter, 1u) : blockIdx.x;
    __syncthreads();
    const int64_t tile = s_tile;
    const int w = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int64_t wbase = tile * S::TILE + (int64_t)w * S::WAVE_ELEMS + (int64_t)lane * V;

    T x[R][V];
    uint32_t hb[R];
    uint32_t eb[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t base = wbase + (int64_t)r * S::ROW;
        load_lane<Op, V>(in, base, n, aligned, x[r]);
        hb[r] = 0;
        eb[r] = 0;
        if constexpr (SEG) {
            if (base < n) {
                const int64_t wd = base >> 5;
                const int sh = (int)(base & 31);
                const uint64_t two = (uint64_t)sb.heads[wd] | ((uint64_t)sb.heads[wd + 1] << 32);
                hb[r] = (uint32_t)(two >> sh) & VMASK;
                if constexpr (REDUCE) {
                    uint32_t nxt = (uint32_t)(two >> (sh + 1)) & VMASK;   // head at element + 1
                    if (base + V >= n) {                                  // the last element ends its segment
                        int64_t last = n - 1 - base;
                        if (last < V) nxt |= 1u << last;
                        nxt &= (last + 1 >= 32) ? 0xffffffffu : ((1u << (last + 1)) - 1u);
                    }
                    eb[r] = nxt;
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
        if constexpr (SEG) seg_wave_scan(af, a, op);
        else a = wave_inclusive_scan(a, op);
        T ea = shfl_up1(a);
        int ef = SEG ? shfl_up1(af) : 0;
        if (lane == 0) { ea = Op::neutral(); ef = 0; }
        const T ra = bcast_last(a);
        const int rf = SEG ? bcast_last(af) : 0;
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
        if (tile == 0) {
            if (lane == 0) St::publish(status, 0, kPrefix, tv);
        } else {
            const bool done = SEG && tf;      // a head inside the tile: later tiles need nothing earlier
            if (lane == 0) St::publish(status, tile, done ? kPrefix : kAggregate, tv);
            bool first_head = false;
            if constexpr (SEG) {
                const int64_t t0 = tile * S::TILE;
                first_head = (sb.heads[t0 >> 5] >> (t0 & 31)) & 1u;
            }
            if (!first_head) P = lookback(status, tile, op, err);
            if (!done && lane == 0) St::publish(status, tile, kPrefix, op(P, tv));
        }
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < S::WAVES; i++) s_win[i] = wpf[i] ? wpv[i] : op(P, wpv[i]);
        }
    }
    __syncthreads();

    const T inc = s_win[w];
    const T cid = Op::cvl_id();
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t base = wbase + (int64_t)r * S::ROW;
        T y[V];
#pragma unroll
        for (int k = 0; k < V; k++) {
            T v = x[r][k];
            if ((needs >> (r * V + k)) & 1ull) v = op(inc, v);
            y[k] = op(cid, v);
        }
        if constexpr (!REDUCE) {
            store_lane<T, V>(out, base, n, aligned, y);
        } else {
#pragma unroll
            for (int k = 0; k < V; k++) {
                if ((eb[r] >> k) & 1u) {
                    const int64_t e = base + k;
                    const uint32_t wd = sb.heads[e >> 5];
                    const uint32_t mask = (uint32_t)(((uint64_t)2 << (e & 31)) - 1);
                    const int rk = sb.wcnt[e >> 5] + __popc(wd & mask);
                    out[sb.nz[rk - 1]] = y[k];
                }
            }
        }
    }
}

// Unsegmented reduction.  Block b folds a fixed contiguous range (the
// partition depends only on n), each wave a contiguous quarter of it, rows
// in order; partials are combined in block order by the last block to
// finish.  The result is therefore deterministic, and a left fold for the
// order-sensitive double max/min.
template <class Op, int R>
__global__ __launch_bounds__(kBlock) void k_reduce(const typename Op::T *__restrict__ in, int64_t n,
                                                   int64_t per_block, typename Op::T *__restrict__ partials,
                                                   unsigned *__restrict__ counter, typename Op::T *__restrict__ result,
                                                   int aligned_i) {
    using T = typename Op::T;
    using S = ScanShape<Op, R>;
    constexpr int V = S::V;
    const Op op;
    const bool aligned = aligned_i != 0;
    __shared__ T s_w[S::WAVES];
    __shared__ bool s_last;

    const int w = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int64_t lo = (int64_t)blockIdx.x * per_block;
    const int64_t hi = lo + per_block < n ? lo + per_block : n;
    const int64_t per_wave = per_block / S::WAVES;    // a multiple of ROW
    const int64_t wlo = lo + w * per_wave;
    const int64_t whi = wlo + per_wave < hi ? wlo + per_wave : hi;

    T acc = Op::neutral();
    constexpr int U = 4;                 // rows in flight per lane
    for (int64_t rb = wlo; rb < whi; rb += U * S::ROW) {
        T x[U][V];
#pragma unroll
        for (int u = 0; u < U; u++) load_lane<Op, V>(in, rb + (int64_t)u * S::ROW + (int64_t)lane * V, whi, aligned, x[u]);
#pragma unroll
        for (int u = 0; u < U; u++) {
            T a = x[u][0];
#pragma unroll
            for (int k = 1; k < V; k++) a = op(a, x[u][k]);
            if constexpr (ordered<Op>::value) acc = op(acc, wave_reduce(a, op));
            else acc = op(acc, a);
        }
    }
    if constexpr (!ordered<Op>::value) acc = wave_reduce(acc, op);
    if (lane == 0) s_w[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T b = s_w[0];
        for (int i = 1; i < S::WAVES; i++) b = op(b, s_w[i]);
        __hip_atomic_store(partials + blockIdx.x, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (counter) {
            unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            s_last = prev == gridDim.x - 1;
        } else {
            s_last = true;                  // a one-block launch finishes alone
        }
    }
    __syncthreads();
    if (s_last && w == 0) {
        T tot = Op::neutral();
        for (int i0 = 0; i0 < (int)gridDim.x; i0 += kWave) {
            const int i = i0 + lane;
            T v = i < (int)gridDim.x ? __hip_atomic_load(partials + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                     : Op::neutral();
            tot = op(tot, wave_reduce(v, op));
        }
        if (lane == 0) result[0] = op(Op::cvl_id(), tot);
    }
}

template <class T>
__global__ void k_set1(T *p, T v) {
    if (threadIdx.x == 0) p[0] = v;
}

// Writes the identity for every empty segment (segmented reductions).
template <class T>
__global__ void k_empty_identity(T *__restrict__ out, const int *__restrict__ lens, int64_t m, T id) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; g < m; g += (int64_t)gridDim.x * blockDim.x)
        if (lens[g] == 0) out[g] = id;
}

// ------------------------------------------------------------------ host side
constexpr int64_t kScanHeaderBytes = 256;   // tile counter + padding

template <class Op>
int64_t scan_scratch_bytes(int64_t n) {
    using S = ScanShape<Op, kScanRows>;
    int64_t tiles = (n + S::TILE - 1) / S::TILE;
    return kScanHeaderBytes + tiles * TileStatus<typename Op::T>::kWords * 8 + 256;
}

static inline bool al16(const void *p) { return ((uintptr_t)p & 15) == 0; }

template <class Op, int R, bool SEG, bool REDUCE>
static void launch_scan_r(const void *s, void *d, int64_t n, SegBits sb, void *scratch) {
    using S = ScanShape<Op, R>;
    if (n <= 0) return;
    if (!scratch) usage_error("scan of %lld elements called without scratch", (long long)n);
    int64_t tiles = (n + S::TILE - 1) / S::TILE;
    if (tiles > 0x7fffffff) usage_error("vector too long");
    // one tile needs neither the counter nor cleared status words
    if (tiles > 1) fill_u32(scratch, 0, (kScanHeaderBytes + tiles * TileStatus<typename Op::T>::kWords * 8) / 4);
    unsigned *counter = tiles > 1 ? (unsigned *)scratch : nullptr;
    uint64_t *status = (uint64_t *)((char *)scratch + kScanHeaderBytes);
    int aligned = al16(s) && al16(d);
    k_scan<Op, R, SEG, REDUCE><<<(unsigned)tiles, kBlock, 0, g_stream>>>(
        (const typename Op::T *)s, (typename Op::T *)d, n, sb, status, counter, aligned, device_error_word());
    LAUNCHCHECK();
}

template <class Op, bool SEG, bool REDUCE>
static void launch_scan(const void *s, void *d, int64_t n, SegBits sb, void *scratch) {
    launch_scan_r<Op, kScanRows, SEG, REDUCE>(s, d, n, sb, scratch);
}

template <class Op>
void scan_unseg(void *d, const void *s, int64_t n, void *scratch) {
    launch_scan<Op, false, false>(s, d, n, SegBits{nullptr, nullptr, nullptr}, scratch);
}

template <class Op>
void scan_seg(void *d, const void *s, const void *sd, int64_t n, int64_t m, void *scratch) {
    if (m == 1) { scan_unseg<Op>(d, s, n, scratch); return; }   // vprims.c:75
    SegdView v(sd, n, m);
    launch_scan<Op, true, false>(s, d, n, SegBits{v.heads, v.wcnt, v.nz}, scratch);
}

template <class Op>
int64_t reduce_scratch_bytes(int64_t n) {
    (void)n;
    return 256 + 1024 * 8 + 256;
}

static void *g_host_result = nullptr;

template <class Op>
void reduce_unseg_dev(void *result_dev, const void *s, int64_t n, void *scratch) {
    using T = typename Op::T;
    using S = ScanShape<Op, kScanRows>;
    if (!scratch) usage_error("reduction called without scratch");
    unsigned *counter = (unsigned *)scratch;
    T *partials = (T *)((char *)scratch + 256);
    if (n <= 0) {
        k_set1<T><<<1, 64, 0, g_stream>>>((T *)result_dev, Op::cvl_id());
        LAUNCHCHECK();
        return;
    }
    int64_t tiles = (n + S::TILE - 1) / S::TILE;
    int grid = (int)(tiles < 1024 ? tiles : 1024);
    int64_t per_block = (n + grid - 1) / grid;
    per_block = (per_block + S::WAVES * S::ROW - 1) / (S::WAVES * S::ROW) * (S::WAVES * S::ROW);
    grid = (int)((n + per_block - 1) / per_block);
    if (grid > 1) fill_u32(counter, 0, 1);
    k_reduce<Op, kScanRows><<<grid, kBlock, 0, g_stream>>>((const T *)s, n, per_block, partials,
                                                           grid > 1 ? counter : nullptr, (T *)result_dev, al16(s));
    LAUNCHCHECK();
}

template <class Op>
typename Op::T reduce_unseg(const void *s, int64_t n, void *scratch) {
    using T = typename Op::T;
    if (!g_host_result) HIPCHECK(hipHostMalloc(&g_host_result, 64, hipHostMallocDefault));
    if (n <= 0) return Op::cvl_id();
    T *res = (T *)((char *)scratch + 256 + 1024 * 8);
    reduce_unseg_dev<Op>(res, s, n, scratch);
    HIPCHECK(hipMemcpyAsync(g_host_result, res, sizeof(T), hipMemcpyDeviceToHost, g_stream));
    HIPCHECK(hipStreamSynchronize(g_stream));
    return *(T *)g_host_result;
}

template <class Op>
void reduce_seg(void *d, const void *s, const void *sd, int64_t n, int64_t m, void *scratch) {
    using T = typename Op::T;
    if (m <= 0) return;
    if (m == 1) {                                    // vprims.c:167
        reduce_unseg_dev<Op>(d, s, n, scratch);
        return;
    }
    SegdView v(sd, n, m);
    int g = grid_for(m, kBlock);
    if (g > 4096) g = 4096;
    k_empty_identity<T><<<g, kBlock, 0, g_stream>>>((T *)d, v.lens, m, Op::cvl_id());
    LAUNCHCHECK();
    launch_scan<Op, true, true>(s, d, n, SegBits{v.heads, v.wcnt, v.nz}, scratch);
}

// The same with 8-row tiles (a schedule the segmented-op dispatcher may
// pick; the scratch it needs is twice that of the default tiles).
template <class Op>
void scan_seg_r8(void *d, const void *s, const void *sd, int64_t n, int64_t m, void *scratch) {
    if (m == 1) { launch_scan_r<Op, 8, false, false>(s, d, n, SegBits{nullptr, nullptr, nullptr}, scratch); return; }
    SegdView v(sd, n, m);
    launch_scan_r<Op, 8, true, false>(s, d, n, SegBits{v.heads, v.wcnt, v.nz}, scratch);
}

template <class Op>
void reduce_seg_r8(void *d, const void *s, const void *sd, int64_t n, int64_t m, void *scratch) {
    using T = typename Op::T;
    if (m <= 0) return;
    if (m == 1) { reduce_unseg_dev<Op>(d, s, n, scratch); return; }
    SegdView v(sd, n, m);
    int g = grid_for(m, kBlock);
    if (g > 4096) g = 4096;
    k_empty_identity<T><<<g, kBlock, 0, g_stream>>>((T *)d, v.lens, m, Op::cvl_id());
    LAUNCHCHECK();
    launch_scan_r<Op, 8, true, true>(s, d, n, SegBits{v.heads, v.wcnt, v.nz}, scratch);
}

#define CVLHIP_INST_SEG(OP)                                                                         \
    template void scan_seg<OP>(void *, const void *, const void *, int64_t, int64_t, void *);      \
    template void reduce_seg<OP>(void *, const void *, const void *, int64_t, int64_t, void *);    \
    template void scan_seg_r8<OP>(void *, const void *, const void *, int64_t, int64_t, void *);   \
    template void reduce_seg_r8<OP>(void *, const void *, const void *, int64_t, int64_t, void *);
CVLHIP_INST_SEG(AddZ)
CVLHIP_INST_SEG(MulZ)
CVLHIP_INST_SEG(MinZ)
CVLHIP_INST_SEG(MaxZ)
CVLHIP_INST_SEG(AddD)
CVLHIP_INST_SEG(MulD)
CVLHIP_INST_SEG(MinD)
CVLHIP_INST_SEG(MaxD)
template int64_t scan_scratch_bytes<AddD>(int64_t);

// Explicit instantiations used by other translation units.
template void scan_unseg<AddZ>(void *, const void *, int64_t, void *);
template void scan_unseg<CountB>(void *, const void *, int64_t, void *);
template void reduce_seg<CountB>(void *, const void *, const void *, int64_t, int64_t, void *);
template CountB::T reduce_unseg<CountB>(const void *, int64_t, void *);
template int64_t scan_scratch_bytes<AddZ>(int64_t);
template AddZ::T reduce_unseg<AddZ>(const void *, int64_t, void *);

}  // namespace cvlhip

using namespace cvlhip;

// ------------------------------------------------------------------ CVL entry points
#define CVLHIP_SCANS(OPN, TY, OP)                                                            \
    void OPN##_su##TY(vec_p d, vec_p s, int len, vec_p scratch) {                            \
        scan_unseg<OP>(d, s, len, scratch);                                                 \
    }                                                                                       \
    int OPN##_su##TY##_scratch(int len) { return units_for_bytes(scan_scratch_bytes<OP>(len)); } \
    unsigned int OPN##_su##TY##_inplace(void) { return INPLACE_1; }                          \
    void OPN##_se##TY(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch) {             \
        scan_seg<OP>(d, s, sd, n, m, scratch);                                              \
    }                                                                                       \
    int OPN##_se##TY##_scratch(int n, int m) {                                               \
        (void)m;                                                                            \
        return units_for_bytes(scan_scratch_bytes<OP>(n));                                  \
    }                                                                                       \
    unsigned int OPN##_se##TY##_inplace(void) { return INPLACE_1; }                          \
    OP::T OPN##_ru##TY(vec_p s, int len, vec_p scratch) { return reduce_unseg<OP>(s, len, scratch); } \
    int OPN##_ru##TY##_scratch(int len) { return units_for_bytes(reduce_scratch_bytes<OP>(len)); } \
    unsigned int OPN##_ru##TY##_inplace(void) { return INPLACE_NONE; }                       \
    void OPN##_re##TY(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch) {             \
        reduce_seg<OP>(d, s, sd, n, m, scratch);                                            \
    }                                                                                       \
    int OPN##_re##TY##_scratch(int n, int m) {                                               \
        (void)m;                                                                            \
        int64_t a = scan_scratch_bytes<OP>(n), b = reduce_scratch_bytes<OP>(n);             \
        return units_for_bytes(a > b ? a : b);                                              \
    }                                                                                       \
    unsigned int OPN##_re##TY##_inplace(void) { return INPLACE_NONE; }

extern "C" {
CVLHIP_SCANS(add, z, AddZ)
CVLHIP_SCANS(add, d, AddD)
CVLHIP_SCANS(mul, z, MulZ)
CVLHIP_SCANS(mul, d, MulD)
CVLHIP_SCANS(min, z, MinZ)
CVLHIP_SCANS(min, d, MinD)
CVLHIP_SCANS(max, z, MaxZ)
CVLHIP_SCANS(max, d, MaxD)
CVLHIP_SCANS(and, b, AndB)
CVLHIP_SCANS(and, z, AndZ)
CVLHIP_SCANS(ior, b, IorB)
CVLHIP_SCANS(ior, z, IorZ)
CVLHIP_SCANS(xor, b, XorB)
CVLHIP_SCANS(xor, z, XorZ)
}
