// pairpass.hpp -- add_sed and max_red of one double vector in one pass
// (cvlhip_add_sed_max_red): the segmented + scan (ref
// nesl/cvl/serial/vprims.c:62-109) and the segmented max reduction
// (vprims.c:155-201) of the same values, reading them once.
//
// The scan is the flat segmented scan of scan_tile (8-byte values, decoupled
// look-back on the sums).  The maximum rides along in the same registers but
// needs no look-back: a segment that ends inside the wave that saw its head
// is complete there and is written at once; the one segment per wave that
// entered it from earlier waves is finished after the tile's wave prefixes
// are known; the one per tile that entered it from earlier tiles is left as
// a record (NolbRecords) for k_seg_reduce_fix.  Rows of a wave with no head
// scan without flags, and only rows with an end compute output slots.
#pragma once

#include "scan_kernel.hpp"

namespace cvlhip {

#ifndef CVLHIP_PAIR_ROWS
#define CVLHIP_PAIR_ROWS 8
#endif
#ifndef CVLHIP_PAIR_BLOCK
#define CVLHIP_PAIR_BLOCK 512
#endif
// Bytes each lane holds per row (one 16-byte load by default).  Wider lanes
// fold their consecutive doubles in order before the row's wave scans, as
// the reduction tiles do, but here every element is also stored: 64-byte
// lanes ran 0.415 ms against 0.362 (2^26 doubles, Zipf(1.5)).
#ifndef CVLHIP_PAIR_LANE_BYTES
#define CVLHIP_PAIR_LANE_BYTES 16
#endif
constexpr int kPairRows = CVLHIP_PAIR_ROWS;
constexpr int kPairBlock = CVLHIP_PAIR_BLOCK;

template <int R, int B>
struct PairShape {
    static constexpr int V = CVLHIP_PAIR_LANE_BYTES / 8;   // doubles per lane and row
    static_assert(V % 2 == 0 && V <= 16, "whole 16-byte loads; end flags in bits 0..15");
    static constexpr int ROW = kWave * V;
    static constexpr int WAVES = B / kWave;
    static constexpr int WAVE_ELEMS = R * ROW;
    static constexpr int64_t TILE = (int64_t)WAVES * WAVE_ELEMS;
};

// Scratch after the 256-byte tile-counter header: the scan's status words
// for ntiles tiles, then the maximum's per-tile records.
inline int64_t pair_status_bytes(int64_t ntiles) {
    return (ntiles * TileStatus<double>::kWords * 8 + 255) & ~(int64_t)255;
}
__host__ __device__ inline NolbRecords<double> pair_records(void *status, int64_t ntiles) {
    return NolbRecords<double>::at((char *)status + ((ntiles * TileStatus<double>::kWords * 8 + 255) & ~(int64_t)255),
                                   ntiles);
}

template <int R, int B>
__global__ __launch_bounds__(B) void k_pair_scan_max(const double *__restrict__ s, double *__restrict__ scan,
                                                     double *__restrict__ red, int64_t n, SegBits sb,
                                                     uint64_t *__restrict__ status, unsigned *__restrict__ counter,
                                                     int aligned_i, unsigned *__restrict__ err) {
    using S = PairShape<R, B>;
    constexpr int V = S::V;
    constexpr uint32_t VMASK = (1u << V) - 1u;
    constexpr int HW = (int)(S::TILE / 32) + 2;
    const AddD add;
    const MaxD mx;
    const double ninf = MaxD::neutral();

    __shared__ unsigned s_tile;
    __shared__ double s_wa[S::WAVES], s_wb[S::WAVES];   // wave aggregates of the sum and the max
    __shared__ int s_wf[S::WAVES];                      // the wave holds a head
    __shared__ double s_ina[S::WAVES], s_inb[S::WAVES]; // what precedes each wave (sum: from the vector's start)
    __shared__ int s_inf[S::WAVES];                     // a head in an earlier wave of the tile
    __shared__ int s_dslot[S::WAVES];                   // the wave's end of the segment entering it (-1: none)
    __shared__ double s_dval[S::WAVES];                 // that segment's max over the wave
    __shared__ int s_fix_slot;
    __shared__ double s_fix_val;
    __shared__ uint32_t s_heads[HW];
    __shared__ int s_wcnt[HW];

    const int tid = threadIdx.x;
    if (tid == 0) {
        // tile ids in launch order (a look-back waits only on started tiles)
        s_tile = counter ? atomicAdd(counter, 1u) : blockIdx.x;
        s_fix_slot = -1;
        s_fix_val = ninf;
    }
    if (tid < S::WAVES) s_dslot[tid] = -1;
    __syncthreads();
    const int64_t tile = s_tile, ntiles = gridDim.x;
    const bool aligned = aligned_i != 0;
    const int w = tid / kWave, lane = tid % kWave;
    const int64_t t0 = tile * S::TILE;
    const int64_t wbase = t0 + (int64_t)w * S::WAVE_ELEMS + (int64_t)lane * V;

    double x[R][V];
#pragma unroll
    for (int r = 0; r < R; r++) load_lane<AddD, V>(s, wbase + (int64_t)r * S::ROW, n, aligned, x[r]);
    // the tile's head bitmap words and head counts, staged once
    const int64_t w0 = t0 >> 5, W = (n + 31) >> 5;
    for (int i = tid; i < HW; i += B) {
        const int64_t wd = w0 + i;
        s_heads[i] = wd <= W ? sb.heads[wd] : 0u;
        s_wcnt[i] = wd <= W ? sb.wcnt[wd] : 0;
    }
    const bool nz_ident = sb.m > 0 && sb.wcnt[W] == sb.m;   // no empty segment: slot = rank - 1
    __syncthreads();

    const double cid_b = MaxD::cvl_id();
    double ca = AddD::neutral(), cb = ninf;   // the wave's running sum and max over its rows so far
    int cf = 0;
    uint64_t needs = 0;                       // bit r*V+k: the element's sum lacks what precedes the wave
#pragma unroll
    for (int r = 0; r < R; r++) {
        // the row's heads (hb), ends (eb bits 0..V-1; bits 16.. the heads of
        // the bitmap word before the lane's part) and the heads before that
        // word (wc, read only where an end needs it), from the staged words
        const int64_t base = wbase + (int64_t)r * S::ROW;
        uint32_t hb = 0, eb = 0;
        int wc = 0;
        if (base < n) {
            const int64_t wd = base >> 5;
            const int sh = (int)(base & 31);
            const uint64_t two = (uint64_t)s_heads[wd - w0] | ((uint64_t)s_heads[wd + 1 - w0] << 32);
            hb = (uint32_t)(two >> sh) & VMASK;
            uint32_t nxt = (uint32_t)(two >> (sh + 1)) & VMASK;   // a head at element + 1
            if (base + V >= n) {                                  // the last element ends its segment
                const int64_t last = n - 1 - base;
                if (last < V) nxt |= 1u << last;
                nxt &= (last + 1 >= 32) ? 0xffffffffu : ((1u << (last + 1)) - 1u);
            }
            eb = nxt | ((uint32_t)__popc((uint32_t)two & ((1u << sh) - 1u)) << 16);
            if (nxt) wc = s_wcnt[wd - w0];
        }
        double a = AddD::neutral(), b = ninf;
        int af = 0;
#pragma unroll
        for (int k = 0; k < V; k++) {
            if ((hb >> k) & 1u) { a = x[r][k]; b = x[r][k]; af = 1; }
            else { a = add(a, x[r][k]); b = mx(b, x[r][k]); }
        }
        if (ballot(hb != 0u)) {
            int fa = af, fb = af;
            seg_wave_scan(fa, a, add);
            seg_wave_scan(fb, b, mx);
            af = fa;
        } else {
            a = wave_inclusive_scan(a, add);
            b = wave_inclusive_scan(b, mx);
        }
        const double ra = bcast_last(a), rb = bcast_last(b);
        const int rf = bcast_last(af);
        double ea = shfl_up1(a), ebx = shfl_up1(b);
        int ef = shfl_up1(af);
        if (lane == 0) { ea = AddD::neutral(); ebx = ninf; ef = 0; }
        double pa = ef ? ea : add(ca, ea);
        double pb = ef ? ebx : mx(cb, ebx);
        int pf = cf | ef;
        const bool ends = ballot((eb & VMASK) != 0u) != 0;
#pragma unroll
        for (int k = 0; k < V; k++) {
            double ya;
            if ((hb >> k) & 1u) { ya = AddD::neutral(); pa = x[r][k]; pb = x[r][k]; pf = 1; }
            else { ya = pa; if (!pf) needs |= 1ull << (r * V + k); pa = add(pa, x[r][k]); pb = mx(pb, x[r][k]); }
            if (ends && ((eb >> k) & 1u)) {
                const int rk = wc + (int)(eb >> 16) + __popc(hb & ((2u << k) - 1u));
                const int slot = nz_ident ? rk - 1 : sb.nz[rk - 1];
                if (pf) {
                    red[slot] = mx(cid_b, pb);        // the segment began in this wave: complete
                } else {
                    s_dslot[w] = slot;                // it entered the wave: finished below
                    s_dval[w] = pb;
                }
            }
            x[r][k] = ya;
        }
        ca = rf ? ra : add(ca, ra);
        cb = rf ? rb : mx(cb, rb);
        cf |= rf;
    }

    if (lane == 0) { s_wa[w] = ca; s_wb[w] = cb; s_wf[w] = cf; }
    __syncthreads();

    if (w == 0) {
        double ta = AddD::neutral(), tb = ninf;
        int tf = 0;
        double wpa[S::WAVES], wpb[S::WAVES];
        int wpf[S::WAVES];
#pragma unroll
        for (int i = 0; i < S::WAVES; i++) {
            wpa[i] = ta;
            wpb[i] = tb;
            wpf[i] = tf;
            ta = s_wf[i] ? s_wa[i] : add(ta, s_wa[i]);
            tb = s_wf[i] ? s_wb[i] : mx(tb, s_wb[i]);
            tf |= s_wf[i];
        }
        using St = TileStatus<double>;
        if (lane == 0) {
            const NolbRecords<double> rec = pair_records(status, ntiles);
            rec.agg[tile] = tb;
            rec.flag[tile] = tf;
        }
        double P = AddD::neutral();
        if (tile == 0) {
            if (lane == 0) St::publish(status, 0, kPrefix, ta);
        } else {
            if (lane == 0) St::publish(status, tile, tf ? kPrefix : kAggregate, ta);
            if (!(s_heads[0] & 1u)) P = lookback(status, tile, add, err);   // t0 is a multiple of 32
            if (!tf && lane == 0) St::publish(status, tile, kPrefix, add(P, ta));
        }
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < S::WAVES; i++) {
                s_ina[i] = wpf[i] ? wpa[i] : add(P, wpa[i]);
                s_inb[i] = wpb[i];
                s_inf[i] = wpf[i];
            }
        }
    }
    __syncthreads();

    const double inc = s_ina[w];
    const double cid_a = AddD::cvl_id();
#pragma unroll
    for (int r = 0; r < R; r++) {
        double y[V];
#pragma unroll
        for (int k = 0; k < V; k++) {
            double v = x[r][k];
            if ((needs >> (r * V + k)) & 1ull) v = add(inc, v);
            y[k] = add(cid_a, v);
        }
        store_lane<double, V>(scan, wbase + (int64_t)r * S::ROW, n, aligned, y);
    }
    // the end, in this wave, of the segment that entered it
    if (lane == 0 && s_dslot[w] >= 0) {
        const double v = mx(s_inb[w], s_dval[w]);
        if (s_inf[w]) {
            red[s_dslot[w]] = mx(cid_b, v);
        } else {                              // it entered the tile: k_seg_reduce_fix adds earlier tiles
            s_fix_slot = s_dslot[w];
            s_fix_val = v;
        }
    }
    __syncthreads();
    if (tid == 0) {
        const NolbRecords<double> rec = pair_records(status, ntiles);
        rec.slot[tile] = s_fix_slot;
        rec.val[tile] = s_fix_val;
    }
}

}  // namespace cvlhip
