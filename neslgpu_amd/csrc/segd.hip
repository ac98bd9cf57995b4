// segd.hip -- segment descriptor construction: mke_fov, len_fos, siz_fos,
// cpy_wus (ref nesl/cvl/serial/facilt.c:106-144, elwise.c:291-298).
//
// mke_fov turns the m lengths into the descriptor laid out in segd.hpp:
// offsets (a scan of the lengths), the list of non-empty segments, the head
// bitmap and the per-word head counts.  It runs once per descriptor; every
// segmented kernel then finds boundaries from the bitmap.  Small
// descriptors are built by one workgroup, larger ones by two single-pass
// look-back kernels.
#include "cvl_hip.hpp"
#include "segd.hpp"
#include "scan.hpp"
#include "lookback.hpp"
#include "mlp.hpp"
#include "../../include/cvl.h"
#include "../../include/cvl_hip_ext.h"

#include <stddef.h>
#include <stdlib.h>

// The trained dispatcher network, generated at build time from
// neslgpu_amd/dispatch_model.npz (sched_model.inc in the build directory).
#if __has_include("sched_model.inc")
#include "sched_model.inc"
#endif

namespace cvlhip {

// ---- the dispatcher's pick
// Parameters of the network that picks a schedule for every descriptor of
// at least kSchedMinN elements (mlp.hpp for the layout).
#ifdef CVLHIP_SCHED_MODEL_PARAMS
__device__ float g_sched_params[kParams] = {CVLHIP_SCHED_MODEL_PARAMS};
static int g_sched_model = 1;
#else
__device__ float g_sched_params[kParams];
static int g_sched_model = 0;
#endif
// CVLHIP_SCHED=0 in the environment starts the library with the
// dispatcher off (every segmented op then runs the flat 8-row schedule).
static int sched_enabled_at_load() {
    const char *e = getenv("CVLHIP_SCHED");
    return !(e && e[0] == '0' && e[1] == 0);
}
static int g_sched_enabled = sched_enabled_at_load();

constexpr int kHist = 2 * kLenClasses;

__device__ __forceinline__ void hist_add(unsigned *h, unsigned len) {
    if (len == 0) return;
    int c = 31 - __clz(len);
    if (c >= kLenClasses) c = kLenClasses - 1;
    atomicAdd(h + c, len);
    atomicAdd(h + kLenClasses + c, 1u);
}

// Waves 0..3 of a workgroup of at least 256 threads each pick the schedule
// of one kind of operation (sched_slot order) from the histogram `hist`;
// thread 0 stores the picks in the header.  Every thread of the workgroup
// must call it.
__device__ void sched_pick_block(SegdHeader *hdr, int64_t n, int64_t m, const unsigned *hist, unsigned *s_pick) {
    // The network's parameters and the histogram are staged in LDS with
    // one coalesced pass first, so the forward passes wait for memory once
    // instead of once per layer.
    __shared__ float s_P[kParams];
    __shared__ unsigned s_h[kHist];
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) s_P[i] = g_sched_params[i];
    if (threadIdx.x < kHist) s_h[threadIdx.x] = hist[threadIdx.x];
    __syncthreads();
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (w < 4) {
        const bool red = (w >> 1) != 0, dbl = (w & 1) != 0;
        const float f = sched_feature(lane < kF ? lane : 0, n, m, s_h, s_h + kLenClasses, dbl, red);
        const int a = mlp_pick_wave(s_P, f, sched_admitted(max_len_class(s_h + kLenClasses), m));
        if (lane == 0) s_pick[w] = (unsigned)a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned p = 0;
        for (int k = 0; k < 4; k++) p |= (s_pick[k] & 15u) << (4 * k);
        hdr->sched = (kSchedMagic << 16) | p;
    }
}

// ---- small descriptors: the whole construction in one workgroup
// (m and the bitmap's word count at most kSmallSegd), instead of a dozen
// launches; descriptors of short vectors are built very often by
// interpreted programs.
constexpr int kSmallThreads = 1024;
constexpr int kSmallPer = 4;
constexpr int kSmallSegd = kSmallThreads * kSmallPer;

// Exclusive scan over the workgroup of one value per thread (wrapping
// unsigned +); every thread must call it.  *total receives the sum.
__device__ unsigned block_exclusive_scan(unsigned x, unsigned *s_wave, unsigned *total) {
    const int tid = threadIdx.x, lane = tid % kWave, w = tid / kWave;
    constexpr int W = kSmallThreads / kWave;
    const unsigned inc = wave_inclusive_scan(x, AddU());
    __syncthreads();                                    // s_wave may still be read from a previous call
    if (lane == kWave - 1) s_wave[w] = inc;
    __syncthreads();
    if (tid < kWave) {
        const unsigned t = tid < W ? s_wave[tid] : 0u;
        const unsigned ti = wave_inclusive_scan(t, AddU());
        if (tid < W) s_wave[tid] = ti - t;
        if (tid == W - 1) *total = ti;
    }
    __syncthreads();
    return s_wave[w] + inc - x;
}

__global__ __launch_bounds__(kSmallThreads) void k_segd_small(const int *__restrict__ lens_in, SegdHeader *hdr,
                                                               int *lens, int *offs, int *nz, int *wcnt,
                                                               uint32_t *heads, int64_t n, int64_t m, int64_t words,
                                                               int pick) {
    __shared__ unsigned s_wave[kSmallThreads / kWave];
    __shared__ unsigned s_total;
    __shared__ uint32_t s_heads[kSmallSegd + 8];
    __shared__ unsigned s_hist[kHist];
    __shared__ unsigned s_pick[4];
    const int tid = threadIdx.x;
    if (tid < kHist) s_hist[tid] = 0u;     // ordered by the barriers of the scans below
    unsigned len[kSmallPer], off[kSmallPer], pos[kSmallPer];
    unsigned lsum = 0, psum = 0;
#pragma unroll
    for (int k = 0; k < kSmallPer; k++) {
        const int64_t g = (int64_t)tid * kSmallPer + k;
        len[k] = g < m ? (unsigned)lens_in[g] : 0u;
        off[k] = lsum;
        pos[k] = psum;
        lsum += len[k];
        psum += (g < m && (int)len[k] > 0) ? 1u : 0u;
    }
    const unsigned lbase = block_exclusive_scan(lsum, s_wave, &s_total);
    const unsigned ltot = s_total;
    const unsigned pbase = block_exclusive_scan(psum, s_wave, &s_total);
    for (int64_t i = tid; i < words + 8; i += kSmallThreads) s_heads[i] = 0u;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSmallPer; k++) {
        const int64_t g = (int64_t)tid * kSmallPer + k;
        if (g < m) {
            const int o = (int)(lbase + off[k]);
            if (lens != lens_in) lens[g] = (int)len[k];
            offs[g] = o;
            if ((int)len[k] > 0) {
                nz[pbase + pos[k]] = (int)g;
                if (o >= 0 && (int64_t)o < n) atomicOr(&s_heads[o >> 5], 1u << (o & 31));
                hist_add(s_hist, len[k]);
            }
        }
    }
    if (tid == 0) {
        offs[m] = (int)ltot;
        hdr->n = n;
        hdr->m = m;
        hdr->stats = kStatsMagic;
    }
    __syncthreads();
    if (tid < kHist) hdr->hist[tid] = s_hist[tid];
    if (pick) sched_pick_block(hdr, n, m, s_hist, s_pick);
    else if (tid == 0) hdr->sched = 0u;
    unsigned hc[kSmallPer], csum = 0;
#pragma unroll
    for (int k = 0; k < kSmallPer; k++) {
        const int64_t wd = (int64_t)tid * kSmallPer + k;
        const uint32_t h = wd < words ? s_heads[wd] : 0u;
        hc[k] = csum;
        csum += (unsigned)__popc(h);
        if (wd < words) heads[wd] = h;
    }
    if (tid < 8) heads[words + tid] = 0u;
    const unsigned cbase = block_exclusive_scan(csum, s_wave, &s_total);
#pragma unroll
    for (int k = 0; k < kSmallPer; k++) {
        const int64_t wd = (int64_t)tid * kSmallPer + k;
        if (wd < words) wcnt[wd] = (int)(cbase + hc[k]);
    }
    if (tid == 0) wcnt[words] = (int)s_total;
}

// ---- larger descriptors: two single-pass look-back kernels
// Pass 1 scans the lengths for (offset, number of non-empty segments
// before) as one packed pair, writes offs, the lengths copy, nz and the head
// bits; pass 2 scans the popcounts of the head words into wcnt.  With the
// clears, four launches instead of a dozen.
struct AddUs : AddU {
    __device__ static unsigned sub(unsigned a, unsigned b) { return a - b; }
};

struct SegPair {                    // (sum of lengths mod 2^32) | (count of non-empty) << 32
    using T = uint64_t;
    __host__ __device__ static T neutral() { return 0; }
    __host__ __device__ static T cvl_id() { return 0; }
    __device__ T operator()(T a, T b) const {
        return (uint64_t)(uint32_t)((uint32_t)a + (uint32_t)b) | (((a >> 32) + (b >> 32)) << 32);
    }
    __device__ static T sub(T a, T b) {
        return (uint64_t)(uint32_t)((uint32_t)a - (uint32_t)b) | (((a >> 32) - (b >> 32)) << 32);
    }
};

template <class Op>
__device__ typename Op::T block_tile_prefix(typename Op::T x, typename Op::T *s_wave, typename Op::T *s_tilep,
                                            uint64_t *status, int64_t tile, unsigned *err,
                                            typename Op::T *excl_in_block) {
    // returns the exclusive prefix of this thread's value over all earlier
    // tiles and threads; *s_tilep receives the tile's inclusive total
    using T = typename Op::T;
    using St = TileStatus<T>;
    const Op op;
    const int tid = threadIdx.x, lane = tid % kWave, w = tid / kWave;
    constexpr int W = kSmallThreads / kWave;
    const T inc = wave_inclusive_scan(x, op);
    if (lane == kWave - 1) s_wave[w] = inc;
    __syncthreads();
    if (w == 0) {
        T tv = Op::neutral();
        T wp[W];
#pragma unroll
        for (int i = 0; i < W; i++) { wp[i] = tv; tv = op(tv, s_wave[i]); }
        T P = Op::neutral();
        if (tile == 0) {
            if (lane == 0) St::publish(status, 0, kPrefix, tv);
        } else {
            if (lane == 0) St::publish(status, tile, kAggregate, tv);
            P = lookback(status, tile, op, err);
            if (lane == 0) St::publish(status, tile, kPrefix, op(P, tv));
        }
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < W; i++) s_wave[i] = op(P, wp[i]);
            *s_tilep = op(P, tv);
        }
    }
    __syncthreads();
    *excl_in_block = Op::sub(inc, x);
    return op(s_wave[w], *excl_in_block);
}

// The header's histogram is cleared before the launch; every tile adds its
// segments' counts to it.
__global__ __launch_bounds__(kSmallThreads) void k_segd_offsets(const int *__restrict__ lens_in, SegdHeader *hdr,
                                                                 int *lens, int *offs, int *nz, uint32_t *heads,
                                                                 int64_t n, int64_t m, uint64_t *status,
                                                                 unsigned *counter, unsigned *err) {
    __shared__ unsigned s_tile;
    __shared__ uint64_t s_wave[kSmallThreads / kWave];
    __shared__ uint64_t s_total;
    __shared__ int s_head[kSmallSegd];          // start of each non-empty segment of the tile, else -1
    __shared__ unsigned s_hist[kHist];
    const int tid = threadIdx.x;
    if (tid < kHist) s_hist[tid] = 0u;
    if (tid == 0) s_tile = atomicAdd(counter, 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t g0 = tile * kSmallSegd + (int64_t)tid * kSmallPer;
    unsigned len[kSmallPer];
    uint64_t loc[kSmallPer], sum = 0;
    const SegPair op;
#pragma unroll
    for (int k = 0; k < kSmallPer; k++) {
        const int64_t g = g0 + k;
        len[k] = g < m ? (unsigned)lens_in[g] : 0u;
        loc[k] = sum;
        sum = op(sum, (uint64_t)len[k] | ((uint64_t)(g < m && (int)len[k] > 0) << 32));
    }
    uint64_t excl;
    const uint64_t P = block_tile_prefix<SegPair>(sum, s_wave, &s_total, status, tile, err, &excl);
    int o4[kSmallPer];
#pragma unroll
    for (int k = 0; k < kSmallPer; k++) {
        const int64_t g = g0 + k;
        const uint64_t q = op(P, loc[k]);
        o4[k] = (int)(uint32_t)q;
        s_head[tid * kSmallPer + k] = (g < m && (int)len[k] > 0 && o4[k] >= 0 && (int64_t)o4[k] < n) ? o4[k] : -1;
        if (g < m && (int)len[k] > 0) {
            nz[q >> 32] = (int)g;
            hist_add(s_hist, len[k]);
        }
    }
    static_assert(kSmallPer == 4, "one 16-byte store per thread");
    typedef int i32x4 __attribute__((ext_vector_type(4)));
    // (a descriptor in a CVL arena starts on an 8-byte unit: check)
    const bool a16 = (((uintptr_t)offs | (uintptr_t)lens) & 15) == 0;
    if (a16 && g0 + kSmallPer <= m) {    // g0 is a multiple of 4
        *(i32x4 *)(offs + g0) = i32x4{o4[0], o4[1], o4[2], o4[3]};
        if (lens != lens_in) *(i32x4 *)(lens + g0) = i32x4{(int)len[0], (int)len[1], (int)len[2], (int)len[3]};
    } else {
#pragma unroll
        for (int k = 0; k < kSmallPer; k++) {
            if (g0 + k < m) {
                offs[g0 + k] = o4[k];
                if (lens != lens_in) lens[g0 + k] = (int)len[k];
            }
        }
    }
    __syncthreads();
    if (tid < kHist && s_hist[tid]) atomicAdd(&hdr->hist[tid], s_hist[tid]);
    // Head bits: wave w takes the tile's segments w*256 .. w*256+255, 64
    // consecutive ones per round, ORs the bits of lanes whose heads share a
    // bitmap word (segment starts increase with the segment index) and
    // lets the first lane of each such run set them with one atomic.
    {
        const int w = tid / kWave, lane = tid % kWave;
#pragma unroll
        for (int k = 0; k < kSmallPer; k++) {
            const int o = s_head[w * kWave * kSmallPer + k * kWave + lane];
            int key = o >= 0 ? (o >> 5) : -1 - lane;       // an empty segment matches no neighbour
            uint32_t v = o >= 0 ? 1u << (o & 31) : 0u;
#pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
                const uint32_t vo = __shfl_down(v, d);
                const int ko = __shfl_down(key, d);
                if (lane + d < kWave && ko == key) v |= vo;
            }
            const int kp = __shfl_up(key, 1);
            if (o >= 0 && (lane == 0 || kp != key)) atomicOr(heads + (o >> 5), v);
        }
    }
    if (tile == (int64_t)gridDim.x - 1 && tid == 0) {
        offs[m] = (int)(uint32_t)s_total;
        hdr->n = n;
        hdr->m = m;
        hdr->stats = kStatsMagic;
    }
}

// With `pick`, the launch has one workgroup more than tiles, and that last
// workgroup picks the schedules from the histogram k_segd_offsets completed,
// beside the tiles instead of after one of them.
__global__ __launch_bounds__(kSmallThreads) void k_segd_wcnt(const uint32_t *__restrict__ heads, int *wcnt,
                                                              int64_t words, uint64_t *status, unsigned *counter,
                                                              unsigned *err, SegdHeader *hdr, int64_t n, int64_t m,
                                                              int pick) {
    __shared__ unsigned s_tile;
    __shared__ unsigned s_wave[kSmallThreads / kWave];
    __shared__ unsigned s_total;
    __shared__ unsigned s_pick[4];
    const int tid = threadIdx.x;
    if (pick && blockIdx.x == gridDim.x - 1) {
        sched_pick_block(hdr, n, m, hdr->hist, s_pick);
        return;
    }
    const int64_t ntiles = (int64_t)gridDim.x - (pick ? 1 : 0);
    if (tid == 0) s_tile = atomicAdd(counter, 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t w0 = tile * kSmallSegd + (int64_t)tid * kSmallPer;
    unsigned loc[kSmallPer], sum = 0;
#pragma unroll
    for (int k = 0; k < kSmallPer; k++) {
        const int64_t wd = w0 + k;
        loc[k] = sum;
        sum += wd < words ? (unsigned)__popc(heads[wd]) : 0u;
    }
    unsigned excl;
    const unsigned P = block_tile_prefix<AddUs>(sum, s_wave, &s_total, status, tile, err, &excl);
#pragma unroll
    for (int k = 0; k < kSmallPer; k++) {
        const int64_t wd = w0 + k;
        if (wd < words) wcnt[wd] = (int)(P + loc[k]);
    }
    if (tile == ntiles - 1 && tid == 0) wcnt[words] = (int)s_total;
}

__global__ void k_segd_empty(SegdHeader *hdr, int *offs, int64_t n) {
    if (threadIdx.x == 0) {
        hdr->n = n;
        hdr->m = 0;
        hdr->stats = kStatsMagic;
        offs[0] = 0;
    }
}

static int grid_cap(int64_t n, int cap = 8192) {
    int g = grid_for(n, kBlock);
    return g > cap ? cap : g;
}

int64_t segd_build_scratch_bytes(int64_t n, int64_t m) {
    SegdLayout L(n, m);
    const int64_t ta = (m + kSmallSegd - 1) / kSmallSegd, tb = (L.words + kSmallSegd - 1) / kSmallSegd;
    return 256 + 16 * (ta + 1) + 16 * (tb + 1) + 256;
}

void segd_build(void *sd, const void *lens_in, int64_t n, int64_t m, void *scratch) {
    SegdLayout L(n, m);
    char *b = (char *)sd;
    SegdHeader *hdr = (SegdHeader *)(b + L.hdr);
    const int pick = g_sched_enabled && g_sched_model && n >= kSchedMinN;
    int *lens = (int *)(b + L.lens);
    int *offs = (int *)(b + L.offs);
    int *nz = (int *)(b + L.nz);
    int *wcnt = (int *)(b + L.wcnt);
    uint32_t *heads = (uint32_t *)(b + L.heads);
    if (!scratch) usage_error("mke_fov called without scratch");

    if (m <= kSmallSegd && L.words <= kSmallSegd) {
        k_segd_small<<<1, kSmallThreads, 0, g_stream>>>((const int *)lens_in, hdr, lens, offs, nz, wcnt, heads, n,
                                                        m, L.words, pick);
        LAUNCHCHECK();
        return;
    }
    const int64_t ta = (m + kSmallSegd - 1) / kSmallSegd, tb = (L.words + kSmallSegd - 1) / kSmallSegd;
    unsigned *counters = (unsigned *)scratch;                          // [0] pass 1, [1] pass 2
    uint64_t *st_a = (uint64_t *)((char *)scratch + 256);
    uint64_t *st_b = st_a + 2 * (ta + 1);
    // scratch counters and status, the bitmap, and the header from its
    // schedule word to the end of the histogram, in one launch
    zero3_u32(scratch, (256 + 16 * (ta + 1) + 16 * (tb + 1)) / 4, heads, L.words + 8, &hdr->sched,
              (int64_t)(sizeof(SegdHeader) - offsetof(SegdHeader, sched)) / 4);
    if (m > 0) {
        k_segd_offsets<<<(unsigned)ta, kSmallThreads, 0, g_stream>>>((const int *)lens_in, hdr, lens, offs, nz, heads,
                                                                     n, m, st_a, counters, device_error_word());
        LAUNCHCHECK();
    } else {                                       // no segments: the header and offs[0] = 0
        k_segd_empty<<<1, 64, 0, g_stream>>>(hdr, offs, n);
        LAUNCHCHECK();
    }
    if (L.words > 0) {
        k_segd_wcnt<<<(unsigned)(tb + (pick ? 1 : 0)), kSmallThreads, 0, g_stream>>>(
            heads, wcnt, L.words, st_b, counters + 1, device_error_word(), hdr, n, m, pick);
        LAUNCHCHECK();
    } else {
        fill_u32(wcnt, 0, 1);
    }
}

}  // namespace cvlhip

using namespace cvlhip;

extern "C" {

// ---- the dispatcher's network in the descriptor build
void cvlhip_sched_set_params(const float *params) {
    if (!params) {                    // back to no network: descriptors get no pick
        g_sched_model = 0;
        return;
    }
    HIPCHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(g_sched_params), params, sizeof(float) * kParams, 0,
                                    hipMemcpyDeviceToDevice, g_stream));
    g_sched_model = 1;
}
void cvlhip_sched_get_params(float *params) {
    HIPCHECK(hipMemcpyFromSymbolAsync(params, HIP_SYMBOL(g_sched_params), sizeof(float) * kParams, 0,
                                      hipMemcpyDeviceToDevice, g_stream));
}
int cvlhip_sched_has_model(void) { return g_sched_model; }
void cvlhip_sched_enable(int on) { g_sched_enabled = on != 0; }
int cvlhip_sched_min_n(void) { return (int)kSchedMinN; }

int siz_fos(int n, int m) { return units_for_bytes(SegdLayout(n, m).bytes); }

// Byte offset of the m segment lengths inside a descriptor, so a host may
// read them in place instead of copying them out with len_fos.
int cvlhip_segd_lengths_offset(void) { return (int)SegdLayout(0, 0).lens; }

void mke_fov(vec_p sd, vec_p l, int n, int m, vec_p scratch) { segd_build(sd, l, n, m, scratch); }
int mke_fov_scratch(int n, int m) { return units_for_bytes(segd_build_scratch_bytes(n, m)); }
// The descriptor is larger than the lengths vector, so it cannot be built
// over it.
unsigned int mke_fov_inplace(void) { return INPLACE_NONE; }

void len_fos(vec_p l, vec_p sd, int n, int m, vec_p scratch) {
    (void)scratch;
    if (m > 0)
        HIPCHECK(hipMemcpyAsync(l, (const char *)sd + SegdLayout(n, m).lens, 4 * (size_t)m, hipMemcpyDeviceToDevice,
                                g_stream));
}
int len_fos_scratch(int n, int m) { (void)n; (void)m; return 0; }
// The lengths sit after the descriptor's header, so writing them over the
// descriptor would overlap the source.
unsigned int len_fos_inplace(void) { return INPLACE_NONE; }

// Copy of a whole descriptor (ref elwise.c:291-298 copies the m lengths,
// which is the whole serial descriptor).
void cpy_wus(vec_p d, vec_p s, int n, int m, vec_p scratch) {
    (void)scratch;
    if (d != s) HIPCHECK(hipMemcpyAsync(d, s, (size_t)SegdLayout(n, m).bytes, hipMemcpyDeviceToDevice, g_stream));
}
int cpy_wus_scratch(int n, int m) { (void)n; (void)m; return 0; }
unsigned int cpy_wus_inplace(void) { return INPLACE_1; }

}  // extern "C"
