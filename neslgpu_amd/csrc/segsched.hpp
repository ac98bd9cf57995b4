// segsched.hpp -- the schedules ("arms") of segmented scans and reductions
// and the kernel that runs the one chosen for a descriptor.
//
// A segmented scan or reduction (ref nesl/cvl/serial/vprims.c:62-201) can
// be scheduled several ways; all of them compute the serial semantics:
//   arm 0  flat, 8 rows per wave   decoupled look-back (scans) or two-step
//                                  records (reductions) over the flat vector;
//                                  the fixed schedule of small vectors.
//                                  Reductions: 2 rows of 64 bytes per lane
//                                  (the same elements per tile, a quarter
//                                  of the wave scans)
//   arm 1  flat, 2 rows per wave   4x the tiles: more workgroups on mid-size
//   arm 2  flat, 1 row per wave    vectors, more look-back per element
//   arm 3  thread per segment      one lane walks a segment
//   arm 4  wave per segment        a wave scans 64 x V elements per step
//   arm 5  workgroup per segment   all waves, V elements per lane per step
// Integers, booleans and min/max give bit-identical results on every arm;
// double + and * differ between arms only in association (each arm is
// deterministic), within the compensated-sum bound.
//
// k_seg_sched holds every arm.  The arm comes from the descriptor's header,
// where mke_fov left the dispatcher network's pick, or from the caller (the
// autotuner times each arm that way).  The grid is resident-sized: flat arms
// loop over tiles (scans claim them in order from a counter, so a tile only
// waits on running tiles; reductions take them strided, as they never
// wait), the per-segment arms loop over segments.  The host therefore needs
// no pick to launch it, and the choice costs no round trip.
#pragma once

#include "scan_kernel.hpp"

namespace cvlhip {

enum : int {
    kArmFlat8 = 0,
    kArmFlat2 = 1,
    kArmFlat1 = 2,
    kArmThread = 3,
    kArmWave = 4,
    kArmBlock = 5,
    kSchedArms = 6
};

// Threads per workgroup: those of the fixed schedule of each kind
// (segmented 4-byte scans CVLHIP_SEG_SCAN_BLOCK4, 8-byte scans
// CVLHIP_SEG_SCAN_BLOCK8, reductions CVLHIP_SEG_REDUCE_BLOCK).
#ifndef CVLHIP_SEG_SCAN_BLOCK4
#define CVLHIP_SEG_SCAN_BLOCK4 512
#endif
#ifndef CVLHIP_SEG_SCAN_BLOCK8
#define CVLHIP_SEG_SCAN_BLOCK8 1024
#endif
template <class Op, bool REDUCE>
__host__ __device__ constexpr int sched_block() {
    return REDUCE ? kSegReduceBlock
                  : (sizeof(typename Op::T) == 4 ? CVLHIP_SEG_SCAN_BLOCK4 : CVLHIP_SEG_SCAN_BLOCK8);
}

template <class Op, bool REDUCE, int R, int VB = 16>
__host__ __device__ constexpr int64_t sched_tile() {
    return ScanShape<Op, R, sched_block<Op, REDUCE>(), VB>::TILE;
}
// The 8-row arm's tile: rows of wide lanes for reductions.
template <class Op, bool REDUCE>
__host__ __device__ constexpr int64_t sched_tile8() {
    return REDUCE ? sched_tile<Op, true, kSegReduceRows, kSegReduceLaneBytes>() : sched_tile<Op, false, kSegScanRows>();
}

// Scratch: a 256-byte header (the tile counter of scans at byte 0, the tile
// count of a flat reduction at byte 8), then look-back status words or
// reduction records for the most tiles any arm uses (1-row tiles).
constexpr int64_t kSchedHeader = 256;

template <class Op, bool REDUCE>
int64_t sched_scratch_bytes(int64_t n) {
    using T = typename Op::T;
    constexpr int64_t tile = sched_tile<Op, REDUCE, 1>();
    const int64_t tiles = (n + tile - 1) / tile;
    const int64_t body = REDUCE ? NolbRecords<T>::bytes(tiles) : tiles * TileStatus<T>::kWords * 8;
    return kSchedHeader + body + 256;
}

// ---------------------------------------------------------------- per-segment arms
template <class Op, bool REDUCE, class In>
__device__ __forceinline__ void sched_thread(const In &in, typename Op::T *__restrict__ out, const SegdView &sv) {
    using T = typename Op::T;
    const Op op;
    const T cid = Op::cvl_id();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < sv.m; g += stride) {
        const int64_t lo = sv.offs[g], hi = lo + sv.lens[g];
        T acc = Op::neutral();
        for (int64_t k = lo; k < hi; k++) {
            T x[1];
            in.template load<Op, 1>(k, hi, false, x);
            if constexpr (!REDUCE) out[k] = op(cid, acc);
            acc = op(acc, x[0]);
        }
        if constexpr (REDUCE) out[g] = op(cid, acc);
    }
}

template <class Op, bool REDUCE, class In>
__device__ __forceinline__ void sched_wave(const In &in, typename Op::T *__restrict__ out, const SegdView &sv) {
    using T = typename Op::T;
    constexpr int V = lane_values<Op>::value;
    constexpr int STEP = kWave * V;
    const Op op;
    const T cid = Op::cvl_id();
    const int lane = threadIdx.x % kWave;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    for (int64_t g = wave; g < sv.m; g += nwaves) {
        const int64_t lo = sv.offs[g], hi = lo + sv.lens[g];
        T carry = Op::neutral();
        for (int64_t c = lo; c < hi; c += STEP) {
            T x[V];
            const int64_t b = c + (int64_t)lane * V;
            in.template load<Op, V>(b, hi, false, x);
            T a = x[0];
#pragma unroll
            for (int k = 1; k < V; k++) a = op(a, x[k]);
            const T inc = wave_inclusive_scan(a, op);
            T ex = shfl_up1(inc);
            if (lane == 0) ex = Op::neutral();
            if constexpr (!REDUCE) {
                T p = op(carry, ex);
#pragma unroll
                for (int k = 0; k < V; k++) {
                    if (b + k < hi) out[b + k] = op(cid, p);
                    p = op(p, x[k]);
                }
            }
            carry = op(carry, bcast_last(inc));
        }
        if constexpr (REDUCE) {
            if (lane == 0) out[g] = op(cid, carry);
        }
    }
}

template <class Op, bool REDUCE, class In, int B>
__device__ __forceinline__ void sched_block_arm(const In &in, typename Op::T *__restrict__ out, const SegdView &sv) {
    using T = typename Op::T;
    constexpr int V = lane_values<Op>::value;
    constexpr int W = B / kWave;
    constexpr int STEP = B * V;
    const Op op;
    const T cid = Op::cvl_id();
    __shared__ T s_w[W];
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    for (int64_t g = blockIdx.x; g < sv.m; g += gridDim.x) {
        const int64_t lo = sv.offs[g], hi = lo + sv.lens[g];
        T carry = Op::neutral();
        for (int64_t c = lo; c < hi; c += STEP) {
            T x[V];
            const int64_t b = c + (int64_t)threadIdx.x * V;
            in.template load<Op, V>(b, hi, false, x);
            T a = x[0];
#pragma unroll
            for (int k = 1; k < V; k++) a = op(a, x[k]);
            const T inc = wave_inclusive_scan(a, op);
            if (lane == kWave - 1) s_w[w] = inc;
            __syncthreads();
            T before = Op::neutral(), all = Op::neutral();
#pragma unroll
            for (int i = 0; i < W; i++) {
                const T t = s_w[i];
                if (i < w) before = op(before, t);
                all = op(all, t);
            }
            if constexpr (!REDUCE) {
                T ex = shfl_up1(inc);
                if (lane == 0) ex = Op::neutral();
                T p = op(op(carry, before), ex);
#pragma unroll
                for (int k = 0; k < V; k++) {
                    if (b + k < hi) out[b + k] = op(cid, p);
                    p = op(p, x[k]);
                }
            }
            carry = op(carry, all);
            __syncthreads();
        }
        if constexpr (REDUCE) {
            if (threadIdx.x == 0) out[g] = op(cid, carry);
        }
    }
}

// ---------------------------------------------------------------- flat arms
// The 8-row arm takes one tile per workgroup (the grid always has at least
// as many workgroups as it has tiles).  The finer arms loop over tiles when
// the grid has fewer workgroups than they have tiles; a loop around the
// 8-row tile would cost it a third more registers and an occupancy step.
template <class Op, int R, bool REDUCE, class In, int B, bool LOOP, int VB = 16>
__device__ __forceinline__ void sched_flat(const In &in, typename Op::T *__restrict__ out, int64_t n, SegBits sb,
                                           void *scratch, bool aligned, unsigned *err) {
    constexpr int64_t TILE = ScanShape<Op, R, B, VB>::TILE;
    const int64_t tiles = (n + TILE - 1) / TILE;
    uint64_t *status = (uint64_t *)((char *)scratch + kSchedHeader);
    if constexpr (REDUCE) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *(int64_t *)((char *)scratch + 8) = tiles;
        if constexpr (LOOP) {
            for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
                scan_tile<Op, R, true, true, In, B, true, VB>(in, out, n, sb, status, t, tiles, aligned, err);
                __syncthreads();
            }
        } else {
            const int64_t t = blockIdx.x;
            if (t < tiles)
                scan_tile<Op, R, true, true, In, B, true, VB>(in, out, n, sb, status, t, tiles, aligned, err);
        }
    } else {
        __shared__ unsigned s_t;
        unsigned *counter = (unsigned *)scratch;
        for (;;) {
            if (threadIdx.x == 0) s_t = atomicAdd(counter, 1u);
            __syncthreads();
            const int64_t t = s_t;
            if (t >= tiles) break;
            scan_tile<Op, R, true, false, In, B, false>(in, out, n, sb, status, t, tiles, aligned, err);
            if constexpr (!LOOP) break;
            __syncthreads();
        }
    }
}

// The arm a descriptor's header names for this kind of operation, or the
// fixed schedule when it names none.
template <class Op, bool REDUCE>
__device__ __forceinline__ int sched_arm_of(const void *sd) {
    const uint32_t sc = ((const SegdHeader *)sd)->sched;
    if ((sc >> 16) != kSchedMagic) return kArmFlat8;
    const int a = (int)((sc >> (4 * sched_slot(REDUCE, (int)sizeof(typename Op::T)))) & 15u);
    return a < kSchedArms ? a : kArmFlat8;
}

// sv: the descriptor's view as the host knows it, or (from_hdr) read here
// from the descriptor `sd` itself.  arm < 0: the descriptor's pick.
template <class Op, bool REDUCE, class In>
__global__ __launch_bounds__((sched_block<Op, REDUCE>())) void k_seg_sched(In in, typename Op::T *__restrict__ out,
                                                                          int64_t n, SegdView sv, const void *sd,
                                                                          int from_hdr, void *scratch, int aligned_i,
                                                                          unsigned *err, int arm) {
    constexpr int B = sched_block<Op, REDUCE>();
    if (from_hdr) sv = SegdView::from_header(sd);
    if (arm < 0) arm = sched_arm_of<Op, REDUCE>(sd);
    const SegBits sb{sv.heads, sv.wcnt, sv.nz, nullptr, sv.m};
    const bool aligned = aligned_i != 0;
    switch (arm) {
    case kArmFlat2: sched_flat<Op, 2, REDUCE, In, B, true>(in, out, n, sb, scratch, aligned, err); return;
    case kArmFlat1: sched_flat<Op, 1, REDUCE, In, B, true>(in, out, n, sb, scratch, aligned, err); return;
    case kArmThread:
    case kArmWave:
    case kArmBlock:
        if (REDUCE && blockIdx.x == 0 && threadIdx.x == 0) *(int64_t *)((char *)scratch + 8) = 0;
        if (arm == kArmThread) sched_thread<Op, REDUCE, In>(in, out, sv);
        else if (arm == kArmWave) sched_wave<Op, REDUCE, In>(in, out, sv);
        else sched_block_arm<Op, REDUCE, In, B>(in, out, sv);
        return;
    default:
        sched_flat<Op, REDUCE ? kSegReduceRows : kSegScanRows, REDUCE, In, B, false,
                   REDUCE ? kSegReduceLaneBytes : 16>(in, out, n, sb, scratch, aligned, err);
        return;
    }
}

// Second step of a flat reduction run by k_seg_sched: the tile count comes
// from the scratch header (0 after a per-segment arm: nothing to do).
template <class Op>
__global__ __launch_bounds__(kFixThreads) void k_seg_reduce_fix_sched(void *scratch, typename Op::T *__restrict__ out) {
    const int64_t g = *(const int64_t *)((const char *)scratch + 8);
    if (g <= 1) return;
    seg_reduce_fix_body<Op>(NolbRecords<typename Op::T>::at((char *)scratch + kSchedHeader, g), g, out);
}

// Workgroups of a k_seg_sched launch: one per tile of the 8-row arm, or
// more, up to one per tile of the 1-row arm while they are all resident at
// once.
template <class Op, bool REDUCE, class In>
inline int sched_grid(int64_t n) {
    static int resident = 0;
    if (!resident) {
        int dev = 0, cus = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_seg_sched<Op, REDUCE, In>,
                                                           sched_block<Op, REDUCE>(), 0);
        resident = (cus > 0 ? cus : 1) * (per > 0 ? per : 1);
    }
    constexpr int64_t tile1 = sched_tile<Op, REDUCE, 1>(), tile8 = sched_tile8<Op, REDUCE>();
    const int64_t t1 = (n + tile1 - 1) / tile1, t8 = (n + tile8 - 1) / tile8;
    int64_t g = t1 < resident ? t1 : resident;
    if (g < t8) g = t8;
    return (int)(g > 0 ? g : 1);
}

// The segmented reduction of the CVL entries and of generated fused
// operations, after the identities of empty segments are written.  Vectors
// shorter than kSchedMinN (their descriptors carry no pick) run the fixed
// schedule one tile per workgroup; longer ones, or a forced arm, run
// k_seg_sched and its second step.  Scratch: sched_scratch_bytes<Op, true>.
template <class Op, class In>
inline void seg_reduce_run(const In &in, typename Op::T *out, int64_t n, const SegdView &sv, const void *sd,
                           bool from_hdr, void *scratch, bool aligned, unsigned *err, int arm, hipStream_t st) {
    using T = typename Op::T;
    if (n <= 0) return;
    if (arm < 0 && n < kSchedMinN) {
        using S = ScanShape<Op, kSegReduceRows, kSegReduceBlock, kSegReduceLaneBytes>;
        const int64_t tiles = (n + S::TILE - 1) / S::TILE;
        void *recs = (char *)scratch + kSchedHeader;
        const SegBits sb = from_hdr ? SegBits{nullptr, nullptr, nullptr, sd, 0}
                                    : SegBits{sv.heads, sv.wcnt, sv.nz, nullptr, sv.m};
        k_scan<Op, kSegReduceRows, true, true, In, kSegReduceBlock, true, kSegReduceLaneBytes>
            <<<(unsigned)tiles, kSegReduceBlock, 0, st>>>(
            in, out, n, sb, (uint64_t *)recs, nullptr, aligned, err);
        if (tiles > 1) k_seg_reduce_fix<Op><<<1, kFixThreads, 0, st>>>(NolbRecords<T>::at(recs, tiles), tiles, out);
        return;
    }
    k_seg_sched<Op, true, In><<<sched_grid<Op, true, In>(n), sched_block<Op, true>(), 0, st>>>(
        in, out, n, sv, sd, from_hdr ? 1 : 0, scratch, aligned ? 1 : 0, err, arm);
    k_seg_reduce_fix_sched<Op><<<1, kFixThreads, 0, st>>>(scratch, out);
}

}  // namespace cvlhip
