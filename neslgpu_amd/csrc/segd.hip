// segd.hip -- segment descriptor construction: mke_fov, len_fos, siz_fos,
// cpy_wus (ref nesl/cvl/serial/facilt.c:106-144, elwise.c:291-298).
//
// mke_fov turns the m lengths into the descriptor laid out in segd.hpp:
// offsets (a device scan of the lengths), the list of non-empty segments,
// the head bitmap and the per-word head counts.  It runs once per
// descriptor; every segmented kernel then finds boundaries from the bitmap.
#include "cvl_hip.hpp"
#include "segd.hpp"
#include "scan.hpp"
#include "../../include/cvl.h"

namespace cvlhip {

// offs[m] = offs[m-1] + lens[m-1]
__global__ void k_close_offsets(int *offs, const int *lens, int64_t m) {
    if (threadIdx.x == 0 && blockIdx.x == 0) offs[m] = m > 0 ? (int)((unsigned)offs[m - 1] + (unsigned)lens[m - 1]) : 0;
}

// For every non-empty segment g: nz[pos[g]] = g and set its head bit.
// pos = exclusive count of non-empty segments before g.
__global__ void k_nz_heads(const int *__restrict__ lens, const int *__restrict__ offs, const int *__restrict__ pos,
                           int *__restrict__ nz, uint32_t *__restrict__ heads, int64_t m, int64_t n) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; g < m; g += (int64_t)gridDim.x * blockDim.x) {
        if (lens[g] > 0) {
            nz[pos[g]] = (int)g;
            int64_t o = offs[g];
            if (o >= 0 && o < n) atomicOr(heads + (o >> 5), 1u << (o & 31));
        }
    }
}

__global__ void k_popcount(const uint32_t *__restrict__ heads, int *__restrict__ cnt, int64_t words) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; w < words; w += (int64_t)gridDim.x * blockDim.x) cnt[w] = __popc(heads[w]);
}

__global__ void k_close_wcnt(int *wcnt, const uint32_t *heads, int64_t words) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        wcnt[words] = words > 0 ? wcnt[words - 1] + __popc(heads[words - 1]) : 0;
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

__global__ __launch_bounds__(kSmallThreads) void k_segd_small(const int *__restrict__ lens_in, int *lens, int *offs,
                                                               int *nz, int *wcnt, uint32_t *heads, int64_t n,
                                                               int64_t m, int64_t words) {
    __shared__ unsigned s_wave[kSmallThreads / kWave];
    __shared__ unsigned s_total;
    __shared__ uint32_t s_heads[kSmallSegd + 8];
    const int tid = threadIdx.x;
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
            }
        }
    }
    if (tid == 0) offs[m] = (int)ltot;
    __syncthreads();
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

static int grid_cap(int64_t n, int cap = 8192) {
    int g = grid_for(n, kBlock);
    return g > cap ? cap : g;
}

int64_t segd_build_scratch_bytes(int64_t n, int64_t m) {
    SegdLayout L(n, m);
    int64_t a = scan_scratch_bytes<AddZ>(m > L.words ? m : L.words);
    return a + 4 * (m + 64);
}

void segd_build(void *sd, const void *lens_in, int64_t n, int64_t m, void *scratch) {
    SegdLayout L(n, m);
    char *b = (char *)sd;
    int *lens = (int *)(b + L.lens);
    int *offs = (int *)(b + L.offs);
    int *nz = (int *)(b + L.nz);
    int *wcnt = (int *)(b + L.wcnt);
    uint32_t *heads = (uint32_t *)(b + L.heads);
    if (!scratch) usage_error("mke_fov called without scratch");
    int *pos = (int *)((char *)scratch + (scan_scratch_bytes<AddZ>(m > L.words ? m : L.words) + 255) / 256 * 256);

    if (m <= kSmallSegd && L.words <= kSmallSegd) {
        k_segd_small<<<1, kSmallThreads, 0, g_stream>>>((const int *)lens_in, lens, offs, nz, wcnt, heads, n, m,
                                                        L.words);
        LAUNCHCHECK();
        return;
    }
    if ((const void *)lens != lens_in && m > 0)
        HIPCHECK(hipMemcpyAsync(lens, lens_in, 4 * (size_t)m, hipMemcpyDeviceToDevice, g_stream));
    scan_unseg<AddZ>(offs, lens, m, scratch);
    k_close_offsets<<<1, 64, 0, g_stream>>>(offs, lens, m);
    LAUNCHCHECK();
    scan_unseg<CountB>(pos, lens, m, scratch);
    fill_u32(heads, 0, L.words + 8);
    if (m > 0) {
        k_nz_heads<<<grid_cap(m), kBlock, 0, g_stream>>>(lens, offs, pos, nz, heads, m, n);
        LAUNCHCHECK();
    }
    if (L.words > 0) {
        k_popcount<<<grid_cap(L.words), kBlock, 0, g_stream>>>(heads, wcnt, L.words);
        LAUNCHCHECK();
        scan_unseg<AddZ>(wcnt, wcnt, L.words, scratch);
    }
    k_close_wcnt<<<1, 64, 0, g_stream>>>(wcnt, heads, L.words);
    LAUNCHCHECK();
}

}  // namespace cvlhip

using namespace cvlhip;

extern "C" {

int siz_fos(int n, int m) { return units_for_bytes(SegdLayout(n, m).bytes); }

void mke_fov(vec_p sd, vec_p l, int n, int m, vec_p scratch) { segd_build(sd, l, n, m, scratch); }
int mke_fov_scratch(int n, int m) { return units_for_bytes(segd_build_scratch_bytes(n, m)); }
// The descriptor is larger than the lengths vector, so it cannot be built
// over it.
unsigned int mke_fov_inplace(void) { return INPLACE_NONE; }

void len_fos(vec_p l, vec_p sd, int n, int m, vec_p scratch) {
    (void)n; (void)scratch;
    if (l != sd && m > 0) HIPCHECK(hipMemcpyAsync(l, sd, 4 * (size_t)m, hipMemcpyDeviceToDevice, g_stream));
}
int len_fos_scratch(int n, int m) { (void)n; (void)m; return 0; }
unsigned int len_fos_inplace(void) { return INPLACE_1; }

// Copy of a whole descriptor (ref elwise.c:291-298 copies the m lengths,
// which is the whole serial descriptor).
void cpy_wus(vec_p d, vec_p s, int n, int m, vec_p scratch) {
    (void)scratch;
    if (d != s) HIPCHECK(hipMemcpyAsync(d, s, (size_t)SegdLayout(n, m).bytes, hipMemcpyDeviceToDevice, g_stream));
}
int cpy_wus_scratch(int n, int m) { (void)n; (void)m; return 0; }
unsigned int cpy_wus_inplace(void) { return INPLACE_1; }

}  // extern "C"
