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

template <class Op, int R, int B = kBlock>
struct ScanShape {
    using T = typename Op::T;
    static constexpr int V = lane_values<Op>::value;   // elements per lane per row
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
    const void *sd;          // when set (and heads null): read the rest from this descriptor's header
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

// A loader that also says where results go when one pass produces two of
// them (io_dual): the kernel then writes the exclusive scan of component a
// and the per-segment reduction of component b itself.
template <class In> struct io_dual { static constexpr bool value = false; };

// Two double results of the same values in one pass: component a runs a
// segmented scan with A, component b a segmented reduction with B.
struct PairD {
    double a, b;
};
template <class A, class B>
struct PairOp {
    using T = PairD;
    __host__ __device__ static T neutral() { return {A::neutral(), B::neutral()}; }
    __host__ __device__ static T cvl_id() { return {A::cvl_id(), B::cvl_id()}; }
    __host__ __device__ T operator()(T x, T y) const { return {A()(x.a, y.a), B()(x.b, y.b)}; }
};
template <class A, class B>
struct seq_combine<PairOp<A, B>> {
    static constexpr bool value = seq_combine<A>::value || seq_combine<B>::value;
};
// Two pairs per lane and row: one 16-byte load of two doubles.
#ifndef CVLHIP_PAIR_V
#define CVLHIP_PAIR_V 2
#endif
template <class A, class B>
struct lane_values<PairOp<A, B>> { static constexpr int value = CVLHIP_PAIR_V; };

// Reads a double vector as pairs (x, x) and stores the scan of a into
// `scan` and the reduction of b into `red` (one value per segment).
struct PairIn {
    const double *p;
    double *scan;
    double *red;
    template <class Op, int V>
    __device__ __forceinline__ void load(int64_t base, int64_t n, bool aligned, PairD (&x)[V]) const {
        if constexpr (V == 2) {
            if (aligned && base + 2 <= n) {
                const typename Vec16<double>::type w = *(const typename Vec16<double>::type *)(p + base);
                const double w0 = w[0], w1 = w[1];
                x[0] = {w0, w0};
                x[1] = {w1, w1};
                return;
            }
        }
#pragma unroll
        for (int k = 0; k < V; k++) {
            if (base + k < n) {
                const double v = p[base + k];
                x[k] = {v, v};
            } else {
                x[k] = Op::neutral();
            }
        }
    }
};
template <> struct io_dual<PairIn> { static constexpr bool value = true; };

// One kernel for: unsegmented exclusive scan (SEG=false, REDUCE=false),
// segmented exclusive scan (SEG=true, REDUCE=false), segmented reduce
// (SEG=true, REDUCE=true; `out` has one element per segment) and, with a
// dual loader (SEG=true, REDUCE=true), both at once on a pair.
template <class Op, int R, bool SEG, bool REDUCE, class In = ArrayIn<typename Op::T>, int B = kBlock>
__global__ __launch_bounds__(B) void k_scan(In in, typename Op::T *__restrict__ out, int64_t n, SegBits sb,
                                            uint64_t *__restrict__ status, unsigned *__restrict__ counter,
                                            int aligned_i, unsigned *__restrict__ err) {
    using T = typename Op::T;
    using S = ScanShape<Op, R, B>;
    using St = TileStatus<T>;
    constexpr int V = S::V;
    constexpr uint32_t VMASK = (1u << V) - 1u;
    constexpr bool DUAL = io_dual<In>::value;
    static_assert(!DUAL || (SEG && REDUCE), "a dual pass is a segmented scan and reduction");
    const Op op;
    const bool aligned = aligned_i != 0;

    __shared__ unsigned s_tile;
    __shared__ T s_wagg[S::WAVES];
    __shared__ int s_wflag[S::WAVES];
    __shared__ T s_win[S::WAVES];

    // tile ids in launch order (look-back waits only on started tiles); a
    // single-tile launch passes no counter
    if (threadIdx.x == 0) s_tile = counter ? atomicAdd(counter, 1u) : blockIdx.x;
    if constexpr (SEG) {
        if (!sb.heads) {
            const SegdView v = SegdView::from_header(sb.sd);
            sb.heads = v.heads;
            sb.wcnt = v.wcnt;
            sb.nz = v.nz;
        }
    }
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
        in.template load<Op, V>(base, n, aligned, x[r]);
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
            if constexpr (DUAL) {             // a: exclusive, b: inclusive; same `needs`
                if (h) { y.a = Op::neutral().a; p = x[r][k]; pf = 1; }
                else { y.a = p.a; if (!pf) needs |= 1ull << (r * V + k); p = op(p, x[r][k]); }
                y.b = p.b;
            } else if constexpr (!REDUCE) {
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
        if constexpr (DUAL) {
#pragma unroll
            for (int k = 0; k < V; k++) {
                const int64_t e = base + k;
                if (e < n) in.scan[e] = y[k].a;
                if ((eb[r] >> k) & 1u) {
                    const uint32_t wd = sb.heads[e >> 5];
                    const uint32_t mask = (uint32_t)(((uint64_t)2 << (e & 31)) - 1);
                    const int rk = sb.wcnt[e >> 5] + __popc(wd & mask);
                    in.red[sb.nz[rk - 1]] = y[k].b;
                }
            }
        } else if constexpr (!REDUCE) {
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
