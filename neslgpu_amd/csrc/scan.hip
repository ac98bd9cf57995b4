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
#include "scan_kernel.hpp"
#include "segsched.hpp"
#include "pairpass.hpp"
#include "../../include/cvl.h"

namespace cvlhip {

// The block partials of a reduction are combined by k_reduce_finish, a
// second launch of one wave, in block order.  Combining them in the last
// block to finish would need an acquire-release counter in every block,
// which on MI355X writes back and invalidates the XCD's L2 each time: at
// 1024 blocks a 1 GiB int32 reduction then reads at 4.8 TB/s instead of
// 6.0 (6.3 at 2048 blocks without it).

// Result of a reduction: fold the g block partials in block order, then
// the CVL identity (as the serial loop starts from it), store it and post
// it to the host when asked.
template <class Op>
__global__ __launch_bounds__(kWave) void k_reduce_finish(const typename Op::T *__restrict__ partials, int g,
                                                         typename Op::T *__restrict__ result, Mailbox *mb,
                                                         unsigned long long seq) {
    using T = typename Op::T;
    const Op op;
    const int lane = threadIdx.x;
    // Every partial is loaded before the fold (they were written on other
    // XCDs: one memory latency instead of one per 64), then folded 64 at a
    // time in block order.
    constexpr int K = kReducePartials / kWave;
    T v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int i = k * kWave + lane;
        v[k] = i < g ? partials[i] : Op::neutral();
    }
    T tot = Op::neutral();
#pragma unroll
    for (int k = 0; k < K; k++)
        if (k * kWave < g) tot = op(tot, wave_reduce(v[k], op));
    if (lane == 0) {
        const T r = op(Op::cvl_id(), tot);
        result[0] = r;
        if (mb) mailbox_post(mb, seq, scalar_bits(r));
    }
}

// A block's partial: to `partials` when the grid has several blocks, else
// (one block holds the whole reduction) straight to the result.
template <class Op>
__device__ __forceinline__ void reduce_block_out(typename Op::T b, typename Op::T *__restrict__ partials,
                                                 typename Op::T *__restrict__ result, Mailbox *mb,
                                                 unsigned long long seq) {
    using T = typename Op::T;
    if (gridDim.x > 1) {
        partials[blockIdx.x] = b;
    } else {
        const T r = Op()(Op::cvl_id(), b);
        result[0] = r;
        if (mb) mailbox_post(mb, seq, scalar_bits(r));
    }
}

// Unsegmented reduction.  Block b folds a fixed contiguous range (the
// partition depends only on n), each wave a contiguous quarter of it, rows
// in order; k_reduce_finish combines the partials in block order.  The
// result is therefore deterministic, and a left fold for the
// order-sensitive double max/min.
template <class Op, int R>
__global__ __launch_bounds__(kBlock) void k_reduce(const typename Op::T *__restrict__ in, int64_t n,
                                                   int64_t per_block, typename Op::T *__restrict__ partials,
                                                   typename Op::T *__restrict__ result, int aligned_i, Mailbox *mb,
                                                   unsigned long long seq) {
    using T = typename Op::T;
    using S = ScanShape<Op, R>;
    constexpr int V = S::V;
    const Op op;
    const bool aligned = aligned_i != 0;
    __shared__ T s_w[S::WAVES];

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
        reduce_block_out<Op>(b, partials, result, mb, seq);
    }
}

template <class T>
__global__ void k_set1(T *p, T v) {
    if (threadIdx.x == 0) p[0] = v;
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

// Threads per scan workgroup.  A workgroup is one tile with one look-back,
// so larger workgroups wait less per element; measured on MI355X with
// 16-row tiles (256 / 512 / 1024 threads): int add scan of 2^28 0.572 /
// 0.556 / 0.507 ms, int segmented add scan 0.526 / 0.506 / 0.570 (the
// 1024-thread figures spill registers; see scan_rows); with 8-row tiles,
// double add scans of 2^27 and of 2^26 in Zipf(1.5) segments run best at
// 1024 (0.537 and 0.317 ms against 0.678 and 0.332 at 512).  Scratch
// sizes assume 256 threads and 16 rows, the most tiles.
#ifndef CVLHIP_SCAN_BLOCK_U4
#define CVLHIP_SCAN_BLOCK_U4 1024
#endif
#ifndef CVLHIP_SCAN_ROWS_U4
#define CVLHIP_SCAN_ROWS_U4 12
#endif
template <class Op, bool SEG, bool REDUCE>
constexpr int scan_block() {
    if constexpr (REDUCE) return kBlock;
    else if constexpr (sizeof(typename Op::T) == 4) return SEG ? 512 : CVLHIP_SCAN_BLOCK_U4;
    else return 1024;
}

// Rows per wave of a scan tile (16 rows in a 1024-thread workgroup spill
// registers).  With the workgroup sizes above, 8 rows against 16: 2^28 int
// add scan 0.502 -> 0.463 ms, segmented 0.512 -> 0.489; 2^27 double add
// scan 0.629 -> 0.537 ms; 2^26 doubles in Zipf(1.5) segments 0.341 ->
// 0.317.  12 rows (128 registers, no spill) take the plain int scan to
// 0.443 ms but slow the others (segmented int 0.485 -> 0.543, double
// 0.536 -> 0.652), so only plain 4-byte scans use them.
template <class Op, bool SEG, bool REDUCE>
constexpr int scan_rows() {
    return REDUCE ? kScanRows : ((sizeof(typename Op::T) == 4 && !SEG) ? CVLHIP_SCAN_ROWS_U4 : 8);
}

// err: the device error word the kernels raise (null: the library's).
template <class Op, int R, bool SEG, bool REDUCE, int B = scan_block<Op, SEG, REDUCE>()>
static void launch_scan_r(const void *s, void *d, int64_t n, SegBits sb, void *scratch, unsigned *err = nullptr) {
    using S = ScanShape<Op, R, B>;
    if (n <= 0) return;
    if (!scratch) { usage_error("scan of %lld elements called without scratch", (long long)n); return; }
    int64_t tiles = (n + S::TILE - 1) / S::TILE;
    if (tiles > 0x7fffffff) { usage_error("vector too long"); return; }
    // one tile needs neither the counter nor cleared status words
    if (tiles > 1) fill_u32(scratch, 0, (kScanHeaderBytes + tiles * TileStatus<typename Op::T>::kWords * 8) / 4);
    unsigned *counter = tiles > 1 ? (unsigned *)scratch : nullptr;
    uint64_t *status = (uint64_t *)((char *)scratch + kScanHeaderBytes);
    int aligned = al16(s) && al16(d);
    k_scan<Op, R, SEG, REDUCE, ArrayIn<typename Op::T>, B><<<(unsigned)tiles, B, 0, g_stream>>>(
        ArrayIn<typename Op::T>{(const typename Op::T *)s}, (typename Op::T *)d, n, sb, status, counter, aligned,
        err ? err : device_error_word());
    LAUNCHCHECK();
}

// The CVL entry points: scratch_bytes counts tiles of kBlock threads and
// kScanRows rows, so the tiles launched must not be smaller.
template <class Op, bool SEG, bool REDUCE>
static void launch_scan(const void *s, void *d, int64_t n, SegBits sb, void *scratch, unsigned *err = nullptr) {
    constexpr int R = scan_rows<Op, SEG, REDUCE>();
    static_assert(scan_block<Op, SEG, REDUCE>() * R >= kBlock * kScanRows, "scratch sizes count the most tiles");
    launch_scan_r<Op, R, SEG, REDUCE>(s, d, n, sb, scratch, err);
}

// Scratch of a segmented reduction of n elements of elem_bytes each under
// any schedule (segsched.hpp).
int64_t seg_reduce_scratch_bytes(int64_t n, int elem_bytes) {
    return elem_bytes == 8 ? sched_scratch_bytes<AddD, true>(n) : sched_scratch_bytes<AddZ, true>(n);
}

template <class Op>
void scan_unseg(void *d, const void *s, int64_t n, void *scratch, unsigned *err) {
    launch_scan<Op, false, false>(s, d, n, SegBits{nullptr, nullptr, nullptr}, scratch, err);
}

// Segmented scan.  Vectors shorter than kSchedMinN (their descriptors carry
// no pick) run the fixed schedule one tile per workgroup; longer ones run
// the schedule the descriptor names (k_seg_sched), or `arm` when one is
// forced.  Scratch: the larger of scan_scratch_bytes and
// sched_scratch_bytes<Op, false>.
template <class Op>
void scan_seg(void *d, const void *s, const void *sd, int64_t n, int64_t m, void *scratch, int arm,
              unsigned *err) {
    using T = typename Op::T;
    if (m == 1) { scan_unseg<Op>(d, s, n, scratch, err); return; }   // vprims.c:75
    SegdView v(sd, n, m);
    if (arm < 0 && n < kSchedMinN) {
        launch_scan<Op, true, false>(s, d, n, SegBits{v.heads, v.wcnt, v.nz}, scratch, err);
        return;
    }
    if (n <= 0) return;
    if (!scratch) { usage_error("segmented scan of %lld elements called without scratch", (long long)n); return; }
    constexpr int64_t tile1 = sched_tile<Op, false, 1>();
    const int64_t tiles1 = (n + tile1 - 1) / tile1;
    fill_u32(scratch, 0, (kSchedHeader + tiles1 * TileStatus<T>::kWords * 8) / 4);
    k_seg_sched<Op, false, ArrayIn<T>><<<sched_grid<Op, false, ArrayIn<T>>(n), sched_block<Op, false>(), 0, g_stream>>>(
        ArrayIn<T>{(const T *)s}, (T *)d, n, v, sd, 0, scratch, al16(s) && al16(d), err ? err : device_error_word(),
        arm);
    LAUNCHCHECK();
}

// Unsegmented reduction for operators whose result does not depend on the
// order of the operands (all but double max/min): wave v of the grid folds
// rows v, v + W, v + 2W, ... (W = waves in the grid), so at any moment the
// whole grid reads one contiguous window of memory, as a plain streaming
// read does.  The block partials are combined in block order, so the
// result is the same on every run.
template <class Op, int U>
__global__ __launch_bounds__(kBlock) void k_reduce_strided(const typename Op::T *__restrict__ in, int64_t n,
                                                           typename Op::T *__restrict__ partials,
                                                           typename Op::T *__restrict__ result, int aligned_i,
                                                           Mailbox *mb, unsigned long long seq) {
    using T = typename Op::T;
    using S = ScanShape<Op, 1>;
    constexpr int V = S::V;
    const Op op;
    const bool aligned = aligned_i != 0;
    __shared__ T s_w[S::WAVES];
    const int w = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int64_t nw = (int64_t)gridDim.x * S::WAVES;
    T acc = Op::neutral();
    for (int64_t r0 = (int64_t)blockIdx.x * S::WAVES + w; r0 * S::ROW < n; r0 += U * nw) {
        T x[U][V];
#pragma unroll
        for (int u = 0; u < U; u++) load_lane<Op, V>(in, (r0 + u * nw) * S::ROW + (int64_t)lane * V, n, aligned, x[u]);
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int k = 0; k < V; k++) acc = op(acc, x[u][k]);
    }
    acc = wave_reduce(acc, op);
    if (lane == 0) s_w[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T b = s_w[0];
        for (int i = 1; i < S::WAVES; i++) b = op(b, s_w[i]);
        reduce_block_out<Op>(b, partials, result, mb, seq);
    }
}

#ifndef CVLHIP_REDUCE_GRID
#define CVLHIP_REDUCE_GRID 2048
#endif
// One row per wave per step.  Unrolling steps (several rows in flight per
// wave, each a grid-width apart) reads slower: 1 GiB of int32 at grid 1024
// reads at 6.2 TB/s with one row per step and 5.0 TB/s with eight.
#ifndef CVLHIP_REDUCE_U
#define CVLHIP_REDUCE_U 1
#endif
static_assert(CVLHIP_REDUCE_GRID <= kReducePartials, "the reduction scratch holds kReducePartials partials");

template <class Op>
int64_t reduce_scratch_bytes(int64_t n) {
    (void)n;
    return kReduceScratchBytes;
}

// The reduction of s into result_dev (device memory); with a mailbox, the
// kernel also posts it to the host under `seq`.
template <class Op>
void reduce_unseg_dev(void *result_dev, const void *s, int64_t n, void *scratch, Mailbox *mb = nullptr,
                      unsigned long long seq = 0) {
    using T = typename Op::T;
    using S = ScanShape<Op, kScanRows>;
    if (!scratch) usage_error("reduction called without scratch");
    T *partials = (T *)((char *)scratch + 256);
    if (n <= 0) {
        k_set1<T><<<1, 64, 0, g_stream>>>((T *)result_dev, Op::cvl_id());
        LAUNCHCHECK();
        return;
    }
    int grid;
    if constexpr (!ordered<Op>::value) {
        using S1 = ScanShape<Op, 1>;
        constexpr int64_t step = (int64_t)S1::WAVES * S1::ROW * CVLHIP_REDUCE_U;   // elements per block per step
        const int64_t want = (n + step - 1) / step;
        grid = (int)(want < CVLHIP_REDUCE_GRID ? want : CVLHIP_REDUCE_GRID);
        k_reduce_strided<Op, CVLHIP_REDUCE_U><<<grid, kBlock, 0, g_stream>>>(
            (const T *)s, n, partials, (T *)result_dev, al16(s), mb, seq);
    } else {
        const int64_t tiles = (n + S::TILE - 1) / S::TILE;
        grid = (int)(tiles < kReducePartials ? tiles : kReducePartials);
        int64_t per_block = (n + grid - 1) / grid;
        per_block = (per_block + S::WAVES * S::ROW - 1) / (S::WAVES * S::ROW) * (S::WAVES * S::ROW);
        grid = (int)((n + per_block - 1) / per_block);
        k_reduce<Op, kScanRows><<<grid, kBlock, 0, g_stream>>>((const T *)s, n, per_block, partials,
                                                               (T *)result_dev, al16(s), mb, seq);
    }
    LAUNCHCHECK();
    if (grid > 1) {
        k_reduce_finish<Op><<<1, kWave, 0, g_stream>>>(partials, grid, (T *)result_dev, mb, seq);
        LAUNCHCHECK();
    }
}

template <class Op>
typename Op::T reduce_unseg(const void *s, int64_t n, void *scratch) {
    using T = typename Op::T;
    if (n <= 0) return Op::cvl_id();
    if (!scratch) usage_error("reduction called without scratch");
    T *res = (T *)((char *)scratch + kReduceResultOffset);
    Mailbox *mb = mailbox_device();
    const unsigned long long seq = mailbox_next();
    reduce_unseg_dev<Op>(res, s, n, scratch, mb, seq);
    return scalar_from_bits<T>(mailbox_wait(seq));
}

// Segmented reduction: identities for the empty segments, then the
// schedule of seg_reduce_run (segsched.hpp): the fixed two-step reduction
// below kSchedMinN elements, else the descriptor's pick or the forced arm.
template <class Op>
void reduce_seg(void *d, const void *s, const void *sd, int64_t n, int64_t m, void *scratch, int arm) {
    using T = typename Op::T;
    if (m <= 0) return;
    if (m == 1) {                                    // vprims.c:167
        reduce_unseg_dev<Op>(d, s, n, scratch);
        return;
    }
    if (!scratch) { usage_error("segmented reduction of %lld elements called without scratch", (long long)n); return; }
    SegdView v(sd, n, m);
    int g = grid_for(m, kBlock);
    if (g > 4096) g = 4096;
    k_empty_identity<T><<<g, kBlock, 0, g_stream>>>((T *)d, v.lens, m, Op::cvl_id());
    LAUNCHCHECK();
    seg_reduce_run<Op, ArrayIn<T>>(ArrayIn<T>{(const T *)s}, (T *)d, n, v, sd, false, scratch, al16(s),
                                   device_error_word(), arm, g_stream);
    LAUNCHCHECK();
}

// add_sed and max_red of the same values in one pass (one read of s):
// k_pair_scan_max (pairpass.hpp), then the maximum's records of segments
// that cross tiles.
int64_t add_max_pair_scratch_bytes(int64_t n) {
    using S = PairShape<kPairRows, kPairBlock>;
    const int64_t tiles = (n + S::TILE - 1) / S::TILE;
    return kScanHeaderBytes + pair_status_bytes(tiles) + NolbRecords<double>::bytes(tiles) + 256;
}

void add_max_pair(double *scan_d, double *red_d, const double *s, const void *sd, int64_t n, int64_t m,
                  void *scratch) {
    using S = PairShape<kPairRows, kPairBlock>;
    if (m <= 0) return;
    if (!scratch) usage_error("add_sed/max_red pair called without scratch");
    SegdView v(sd, n, m);
    int g = grid_for(m, kBlock);
    if (g > 4096) g = 4096;
    k_empty_identity<double><<<g, kBlock, 0, g_stream>>>(red_d, v.lens, m, MaxD::cvl_id());
    LAUNCHCHECK();
    if (n <= 0) return;
    const int64_t tiles = (n + S::TILE - 1) / S::TILE;
    if (tiles > 0x7fffffff) usage_error("vector too long");
    // the scan's status words are cleared (the records are all written)
    if (tiles > 1) fill_u32(scratch, 0, (kScanHeaderBytes + pair_status_bytes(tiles)) / 4);
    unsigned *counter = tiles > 1 ? (unsigned *)scratch : nullptr;
    uint64_t *status = (uint64_t *)((char *)scratch + kScanHeaderBytes);
    k_pair_scan_max<kPairRows, kPairBlock><<<(unsigned)tiles, kPairBlock, 0, g_stream>>>(
        s, scan_d, red_d, n, SegBits{v.heads, v.wcnt, v.nz, nullptr, m}, status, counter,
        al16(s) && al16(scan_d), device_error_word());
    LAUNCHCHECK();
    if (tiles > 1) {
        k_seg_reduce_fix<MaxD><<<1, kFixThreads, 0, g_stream>>>(pair_records(status, tiles), tiles, red_d);
        LAUNCHCHECK();
    }
}

#define CVLHIP_INST_SEG(OP)                                                                         \
    template void scan_seg<OP>(void *, const void *, const void *, int64_t, int64_t, void *, int, unsigned *); \
    template void reduce_seg<OP>(void *, const void *, const void *, int64_t, int64_t, void *, int);
CVLHIP_INST_SEG(AddZ)
CVLHIP_INST_SEG(MulZ)
CVLHIP_INST_SEG(MinZ)
CVLHIP_INST_SEG(MaxZ)
CVLHIP_INST_SEG(AddD)
CVLHIP_INST_SEG(MulD)
CVLHIP_INST_SEG(MinD)
CVLHIP_INST_SEG(MaxD)
CVLHIP_INST_SEG(CountB)
template int64_t scan_scratch_bytes<AddD>(int64_t);

// Explicit instantiations used by other translation units.
template void scan_unseg<AddZ>(void *, const void *, int64_t, void *, unsigned *);
template void scan_unseg<CountB>(void *, const void *, int64_t, void *, unsigned *);
template CountB::T reduce_unseg<CountB>(const void *, int64_t, void *);
template int64_t scan_scratch_bytes<AddZ>(int64_t);
template AddZ::T reduce_unseg<AddZ>(const void *, int64_t, void *);

// ------------------------------------------------------------------ no-allocation entry
// add_sez for callers that cannot allocate (code under HIP graph capture,
// serving loops, hosts that forbid hipMalloc): its scratch -- the look-back
// tile counter and status words, and a device error word -- is `scratch`, a
// static device array each GPU holds until the process ends.  Bytes
// [0, 256) hold the error word, the scan's scratch follows.  Every call on
// a GPU uses that GPU's one array, so calls that overlap on one GPU race.
// The tile counter and status words are cleared by a fill kernel at the
// start of each call; the error word starts at zero (static storage), is
// raised only by a look-back that gave up, and is read and cleared by
// cvlhip_check_device_errors.
constexpr int64_t kNoAllocBytes = 4 << 20;
constexpr int64_t kNoAllocErr = 256;        // bytes before the scan's scratch
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocDevices = 64;
static void *g_noalloc_ptr[kNoAllocDevices];   // device address of `scratch`, per device

// The address of this GPU's `scratch` (looked up on the first call on each
// device, with that device current).  The host-side name of the array is
// not a device address; only this pointer is passed to kernels.
static void *noalloc_scratch() {
    int dev = -1;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices) {
        usage_error("cvlhip_add_sez_noalloc: device %d is beyond the %d devices it serves", dev, kNoAllocDevices);
        return nullptr;
    }
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
            return nullptr;
        }
        g_noalloc_ptr[dev] = p;
    }
    return g_noalloc_ptr[dev];
}

unsigned *noalloc_error_word() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocDevices) return nullptr;
    return (unsigned *)g_noalloc_ptr[dev];
}

// Bytes of `scratch` add_sez needs for n elements.
static int64_t noalloc_need(int64_t n) {
    const int64_t a = scan_scratch_bytes<AddZ>(n), b = sched_scratch_bytes<AddZ, false>(n);
    return kNoAllocErr + (a > b ? a : b);
}

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
        scan_seg<OP>(d, s, sd, n, m, scratch, -1);                                          \
    }                                                                                       \
    int OPN##_se##TY##_scratch(int n, int m) {                                               \
        (void)m;                                                                            \
        const int64_t a = scan_scratch_bytes<OP>(n), b = sched_scratch_bytes<OP, false>(n); \
        return units_for_bytes(a > b ? a : b);                                              \
    }                                                                                       \
    unsigned int OPN##_se##TY##_inplace(void) { return INPLACE_1; }                          \
    OP::T OPN##_ru##TY(vec_p s, int len, vec_p scratch) { return reduce_unseg<OP>(s, len, scratch); } \
    int OPN##_ru##TY##_scratch(int len) { return units_for_bytes(reduce_scratch_bytes<OP>(len)); } \
    unsigned int OPN##_ru##TY##_inplace(void) { return INPLACE_NONE; }                       \
    void OPN##_re##TY(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch) {             \
        reduce_seg<OP>(d, s, sd, n, m, scratch, -1);                                        \
    }                                                                                       \
    int OPN##_re##TY##_scratch(int n, int m) {                                               \
        (void)m;                                                                            \
        const int64_t a = sched_scratch_bytes<OP, true>(n), b = reduce_scratch_bytes<OP>(n); \
        return units_for_bytes(a > b ? a : b);                                              \
    }                                                                                       \
    unsigned int OPN##_re##TY##_inplace(void) { return INPLACE_NONE; }

extern "C" {
// add_sez without a scratch argument, allocating nothing (see above).  The
// scratch of n elements is 768 + 8 ceil(n / 2048) bytes (1-row look-back
// tiles of the dispatched schedule), so n may be at most 1,073,545,216.
void cvlhip_add_sez_noalloc(vec_p d, vec_p s, vec_p sd, int n, int m) {
    if (noalloc_need(n) > kNoAllocBytes) {
        usage_error("cvlhip_add_sez_noalloc: %d elements need %lld bytes of scratch, more than its %lld", n,
                    (long long)noalloc_need(n), (long long)kNoAllocBytes);
        return;
    }
    char *p = (char *)noalloc_scratch();
    if (!p) return;
    scan_seg<AddZ>(d, s, sd, n, m, p + kNoAllocErr, -1, (unsigned *)p);
}

void cvlhip_add_sed_max_red(vec_p scan_d, vec_p red_d, vec_p s, vec_p sd, int n, int m, vec_p scratch) {
    add_max_pair((double *)scan_d, (double *)red_d, (const double *)s, sd, n, m, scratch);
}
int cvlhip_add_sed_max_red_scratch(int n, int m) {
    (void)m;
    return units_for_bytes(add_max_pair_scratch_bytes(n));
}

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
