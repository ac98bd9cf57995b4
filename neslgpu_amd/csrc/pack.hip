// pack.hip -- pack: stable compaction of the elements whose flag is true
// (ref nesl/cvl/serial/library.c:51-164; the VCODE PACK driver is
// nesl/vcode/main.c:577-642).
//
// pk2 is one single-pass kernel: each tile loads its flags and values,
// ranks the true flags with a wave-wide scan of per-lane counts, gets the
// number of true flags before the tile from the decoupled look-back, and
// writes the kept values to their final positions.  Input is read once and
// the output written once.  The segmented pack keeps the flat order
// (library.c:139-164), so it is the same kernel; pk1_lev, the per-segment
// counts, is a segmented reduction.
#include "cvl_hip.hpp"
#include "lookback.hpp"
#include "scan.hpp"
#include "../../include/cvl.h"

namespace cvlhip {

template <class T> struct VecN;
template <> struct VecN<int> {
    static constexpr int V = 4;
    typedef int vt __attribute__((ext_vector_type(4)));
    typedef int ft __attribute__((ext_vector_type(4)));
};
template <> struct VecN<double> {
    static constexpr int V = 2;
    typedef double vt __attribute__((ext_vector_type(2)));
    typedef int ft __attribute__((ext_vector_type(2)));
};

#ifndef CVLHIP_PACK_ROWS
#define CVLHIP_PACK_ROWS 8
#endif
constexpr int kPackRows = CVLHIP_PACK_ROWS;

template <class T, int R, int B = kBlock>
struct PackShape {
    static constexpr int V = VecN<T>::V;
    static constexpr int ROW = kWave * V;
    static constexpr int WAVES = B / kWave;
    static constexpr int WAVE_ELEMS = R * ROW;
    static constexpr int TILE = WAVES * WAVE_ELEMS;
};

template <class T, int R, int B>
__global__ __launch_bounds__(B) void k_pack(const T *__restrict__ s, const int *__restrict__ f,
                                                 T *__restrict__ d, int64_t n, int64_t dlim,
                                                 uint64_t *__restrict__ status, unsigned *__restrict__ counter,
                                                 int aligned_i, unsigned *__restrict__ err) {
    using S = PackShape<T, R, B>;
    constexpr int V = S::V;
    const AddZ add;
    const bool aligned = aligned_i != 0;
    __shared__ unsigned s_tile;
    __shared__ int s_wcnt[S::WAVES];
    __shared__ int s_wbase[S::WAVES];
    __shared__ T s_run[S::WAVES][S::ROW];   // one row's kept values, compacted

    if (threadIdx.x == 0) s_tile = counter ? atomicAdd(counter, 1u) : blockIdx.x;   // no counter: one tile
    __syncthreads();
    const int64_t tile = s_tile;
    const int w = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int64_t wbase = tile * S::TILE + (int64_t)w * S::WAVE_ELEMS + (int64_t)lane * V;

    T x[R][V];
    uint32_t fb[R];
    int pos[R];          // rank of the lane's first kept element within the wave
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t base = wbase + (int64_t)r * S::ROW;
        fb[r] = 0;
        if (aligned && base + V <= n) {
            typename VecN<T>::vt v = *(const typename VecN<T>::vt *)(s + base);
            typename VecN<T>::ft g = *(const typename VecN<T>::ft *)(f + base);
#pragma unroll
            for (int k = 0; k < V; k++) { x[r][k] = v[k]; fb[r] |= (g[k] != 0) << k; }
        } else {
#pragma unroll
            for (int k = 0; k < V; k++) {
                bool in = base + k < n;
                x[r][k] = in ? s[base + k] : T(0);
                fb[r] |= (in && f[base + k] != 0) << k;
            }
        }
    }
    int carry = 0;
    int rtot[R];         // kept values per row of the wave
#pragma unroll
    for (int r = 0; r < R; r++) {
        int c = __popc(fb[r]);
        int inc = wave_inclusive_scan(c, add);
        pos[r] = carry + inc - c;
        rtot[r] = bcast_last(inc);
        carry += rtot[r];
    }
    if (lane == 0) s_wcnt[w] = carry;
    __syncthreads();
    if (w == 0) {
        int tot = 0, wb[S::WAVES];
#pragma unroll
        for (int i = 0; i < S::WAVES; i++) { wb[i] = tot; tot += s_wcnt[i]; }
        int P = 0;
        if (tile == 0) {
            if (lane == 0) TileStatus<int>::publish(status, 0, kPrefix, tot);
        } else {
            if (lane == 0) TileStatus<int>::publish(status, tile, kAggregate, tot);
            P = lookback(status, tile, add, err);
            if (lane == 0) TileStatus<int>::publish(status, tile, kPrefix, P + tot);
        }
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < S::WAVES; i++) s_wbase[i] = P + wb[i];
        }
    }
    __syncthreads();
    // Row by row, the wave compacts its kept values in LDS and writes them
    // as one contiguous run: every store instruction covers 64 consecutive
    // output elements.
    const int64_t ob = s_wbase[w];
    T *run = s_run[w];
    int rbase = 0;            // kept values of earlier rows of this wave
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int rcount = rtot[r];
        int o = pos[r] - rbase;
#pragma unroll
        for (int k = 0; k < V; k++) {
            if ((fb[r] >> k) & 1u) run[o++] = x[r][k];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int i = lane; i < rcount; i += kWave) {
            const int64_t at = ob + rbase + i;
            if (at < dlim) d[at] = run[i];
            else *err = 1u;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        rbase += rcount;
    }
}

constexpr int64_t kPackHeader = 256;

template <class T>
int64_t pack_scratch_bytes(int64_t n) {
    int64_t tiles = (n + PackShape<T, kPackRows>::TILE - 1) / PackShape<T, kPackRows>::TILE;
    return kPackHeader + 8 * tiles + 256;
}

// Threads per pack workgroup (one tile, one look-back).  Measured on MI355X
// with 8-row tiles and 256 / 512 / 1024 threads: 2^28 int 0.735 / 0.599 /
// 0.558 ms, 2^27 double 0.462 / 0.458 / 0.459 (16-row tiles of 256
// threads: 0.694 and 0.578).  Scratch sizes assume 256 threads, the most
// tiles.
template <class T>
constexpr int pack_block() { return 1024; }

// Packs the n elements of s whose flag is set into d, which holds dlim.
template <class T>
void pack(void *d, const void *s, const void *f, int64_t n, int64_t dlim, void *scratch) {
    constexpr int B = pack_block<T>();
    static_assert(B >= kBlock, "scratch sizes count tiles of kBlock threads");
    using S = PackShape<T, kPackRows, B>;
    if (n <= 0) return;
    if (!scratch) usage_error("pack called without scratch");
    int64_t tiles = (n + S::TILE - 1) / S::TILE;
    if (tiles > 1) fill_u32(scratch, 0, (kPackHeader + 8 * tiles) / 4);
    int aligned = (((uintptr_t)s | (uintptr_t)f) & 15) == 0;
    k_pack<T, kPackRows, B><<<(unsigned)tiles, B, 0, g_stream>>>(
        (const T *)s, (const int *)f, (T *)d, n, dlim, (uint64_t *)((char *)scratch + kPackHeader),
        tiles > 1 ? (unsigned *)scratch : nullptr, aligned, device_error_word());
    LAUNCHCHECK();
}

}  // namespace cvlhip

using namespace cvlhip;

extern "C" {

int pk1_luv(vec_p f, int len, vec_p scratch) { return reduce_unseg<CountB>(f, len, scratch); }
int pk1_luv_scratch(int len) { return units_for_bytes(kReduceScratchBytes); (void)len; }
unsigned int pk1_luv_inplace(void) { return INPLACE_NONE; }

void pk1_lev(vec_p ds, vec_p f, vec_p sd, int n, int m, vec_p scratch) {
    reduce_seg<CountB>(ds, f, sd, n, m, scratch);
}
int pk1_lev_scratch(int n, int m) {
    (void)m;
    int64_t a = scan_scratch_bytes<AddZ>(n), b = kReduceScratchBytes;
    const int64_t c = seg_reduce_scratch_bytes(n, 4);       // pk1_lev: the segmented count
    if (c > a) a = c;
    return units_for_bytes(a > b ? a : b);
}
unsigned int pk1_lev_inplace(void) { return INPLACE_NONE; }

#define PK2U(NAME, T)                                                                  \
    void NAME(vec_p d, vec_p s, vec_p f, int src_len, int dest_len, vec_p scratch) {   \
        pack<T>(d, s, f, src_len, dest_len, scratch);                                  \
    }                                                                                  \
    int NAME##_scratch(int len) { return units_for_bytes(pack_scratch_bytes<T>(len)); } \
    unsigned int NAME##_inplace(void) { return INPLACE_NONE; }
PK2U(pk2_luz, int)
PK2U(pk2_lub, int)
PK2U(pk2_lud, double)

#define PK2E(NAME, T)                                                                  \
    void NAME(vec_p d, vec_p s, vec_p f, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, \
              vec_p scratch) {                                                         \
        (void)sd_s; (void)m_s; (void)sd_d; (void)m_d;                                  \
        pack<T>(d, s, f, n_s, n_d, scratch);                                           \
    }                                                                                  \
    int NAME##_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len) {   \
        (void)s_seg_len; (void)d_vec_len; (void)d_seg_len;                             \
        return units_for_bytes(pack_scratch_bytes<T>(s_vec_len));                      \
    }                                                                                  \
    unsigned int NAME##_inplace(void) { return INPLACE_NONE; }
PK2E(pk2_lez, int)
PK2E(pk2_leb, int)
PK2E(pk2_led, double)

}  // extern "C"
