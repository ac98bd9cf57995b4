// permute.hip -- permute, back-permute and their flagged / default variants
// (ref nesl/cvl/serial/vprims.c:347-467, vecops.c:11-286).
//
//   smp  d[i[k]] = s[k]                       (a permutation)
//   bck  d[k]    = s[i[k]]                    (gather)
//   fpm  d[i[k]] = s[k] where f[k]
//   bfp  d[k]    = f[k] ? s[i[k]] : 0
//   dpe  d = v, then d[i[k]] = s[k]
//   dfp  d = v, then d[i[k]] = s[k] where f[k]
// Segmented forms offset each index by the start of its segment (the source
// segment for scatters, the destination segment for gathers), found in O(1)
// from the descriptor's head bitmap.
//
// Indices are checked against the destination (scatter) or source (gather)
// length: an out-of-range index skips the access and raises the device
// error word instead of faulting.  Scatters whose indices may repeat (fpm,
// dpe, dfp) resolve collisions as the serial loop does: the last writer in
// element order wins.  They first record, per destination, the largest
// source position aimed at it (atomicMax), then only that position writes.
#include "cvl_hip.hpp"
#include "segd.hpp"
#include "partition.hpp"
#include "../../include/cvl.h"

namespace cvlhip {

static int pgrid(int64_t n) {
    int g = grid_for(n, kBlock * 4);
    return g > 16384 ? 16384 : g;
}

#define GRID_STRIDE(i, n)                                                        \
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
         i += (int64_t)gridDim.x * blockDim.x)

// Segment base of element k for a scatter / gather; sv.m <= 1 means none.
struct SegBase {
    SegdView key;          // descriptor that owns element k (selects g)
    const int *base_offs;  // offsets of the other side (null: same as key)
    bool seg;
    __device__ __forceinline__ int64_t base(int64_t k) const {
        if (!seg) return 0;
        int g = key.seg_of(k);
        return base_offs ? base_offs[g] : key.offs[g];
    }
};

template <class T>
__global__ void k_scatter(T *__restrict__ d, const T *__restrict__ s, const int *__restrict__ idx,
                          const int *__restrict__ f, int64_t n, int64_t n_d, SegBase sb, unsigned *err) {
    GRID_STRIDE(k, n) {
        if (f && !f[k]) continue;
        int64_t p = sb.base(k) + idx[k];
        if (p >= 0 && p < n_d) d[p] = s[k];
        else *err = 1u;
    }
}

template <class T>
__global__ void k_gather(T *__restrict__ d, const T *__restrict__ s, const int *__restrict__ idx,
                         const int *__restrict__ f, int64_t n_d, int64_t n_s, SegBase sb, unsigned *err) {
    GRID_STRIDE(k, n_d) {
        if (f && !f[k]) { d[k] = T(0); continue; }
        int64_t p = sb.base(k) + idx[k];
        if (p >= 0 && p < n_s) d[k] = s[p];
        else { d[k] = T(0); *err = 1u; }
    }
}

// Last-writer-wins scatter, pass 1: win[p] = max source position aimed at p.
__global__ void k_claim(int *__restrict__ win, const int *__restrict__ idx, const int *__restrict__ f, int64_t n,
                        int64_t n_d, SegBase sb) {
    GRID_STRIDE(k, n) {
        if (f && !f[k]) continue;
        int64_t p = sb.base(k) + idx[k];
        if (p >= 0 && p < n_d) atomicMax(win + p, (int)k);
    }
}

template <class T>
__global__ void k_scatter_won(T *__restrict__ d, const T *__restrict__ s, const int *__restrict__ idx,
                              const int *__restrict__ f, const int *__restrict__ win, int64_t n, int64_t n_d,
                              SegBase sb, unsigned *err) {
    GRID_STRIDE(k, n) {
        if (f && !f[k]) continue;
        int64_t p = sb.base(k) + idx[k];
        if (p >= 0 && p < n_d) {
            if (win[p] == (int)k) d[p] = s[k];
        } else {
            *err = 1u;
        }
    }
}

static SegBase no_seg() {
    SegBase b{SegdView(nullptr, 0, 0), nullptr, false};
    return b;
}

template <class T>
void scatter(void *d, const void *s, const void *idx, const void *f, int64_t n, int64_t n_d, SegBase sb) {
    if (n <= 0) return;
    k_scatter<T><<<pgrid(n), kBlock, 0, g_stream>>>((T *)d, (const T *)s, (const int *)idx, (const int *)f, n, n_d,
                                                    sb, device_error_word());
    LAUNCHCHECK();
}

template <class T>
void gather(void *d, const void *s, const void *idx, const void *f, int64_t n_d, int64_t n_s, SegBase sb) {
    if (n_d <= 0) return;
    k_gather<T><<<pgrid(n_d), kBlock, 0, g_stream>>>((T *)d, (const T *)s, (const int *)idx, (const int *)f, n_d,
                                                     n_s, sb, device_error_word());
    LAUNCHCHECK();
}

// Scatter in which several k may aim at one destination; `win` is n_d ints
// of scratch.
template <class T>
void scatter_lww(void *d, const void *s, const void *idx, const void *f, int64_t n, int64_t n_d, SegBase sb,
                 void *win) {
    if (n <= 0 || n_d <= 0) return;
    if (!win) usage_error("permute with collisions needs scratch");
    fill_u32(win, 0xffffffffu, n_d);
    k_claim<<<pgrid(n), kBlock, 0, g_stream>>>((int *)win, (const int *)idx, (const int *)f, n, n_d, sb);
    LAUNCHCHECK();
    k_scatter_won<T><<<pgrid(n), kBlock, 0, g_stream>>>((T *)d, (const T *)s, (const int *)idx, (const int *)f,
                                                        (const int *)win, n, n_d, sb, device_error_word());
    LAUNCHCHECK();
}

static void copy_default(void *d, const void *v, int64_t n_d, int w) {
    if (d != v && n_d > 0)
        HIPCHECK(hipMemcpyAsync(d, v, (size_t)n_d * w, hipMemcpyDeviceToDevice, g_stream));
}

}  // namespace cvlhip

using namespace cvlhip;

static int lww_units(int n_d) { return units_for_bytes(4 * (int64_t)(n_d > 0 ? n_d : 0) + 256); }

#define SEGB(sd, n, m, other) SegBase{SegdView(sd, n, m), other, (m) > 1}

#define PERMS(TY, T, W)                                                                                   \
    void smp_pu##TY(vec_p d, vec_p s, vec_p i, int len, vec_p scratch) {                                   \
        partition_permute<T>(d, s, (const int *)i, nullptr, len, 0, scratch);                             \
    }                                                                                                     \
    int smp_pu##TY##_scratch(int len) { return units_for_bytes(partition_permute_scratch_bytes(len, W)); } \
    unsigned int smp_pu##TY##_inplace(void) { return INPLACE_NONE; }                                      \
    void smp_pe##TY(vec_p d, vec_p s, vec_p i, vec_p sd, int n, int m, vec_p scratch) {                    \
        partition_permute<T>(d, s, (const int *)i, sd, n, m, scratch);                                    \
    }                                                                                                     \
    int smp_pe##TY##_scratch(int n, int m) { (void)m; return units_for_bytes(partition_permute_scratch_bytes(n, W)); } \
    unsigned int smp_pe##TY##_inplace(void) { return INPLACE_NONE; }                                      \
    void bck_pu##TY(vec_p d, vec_p s, vec_p i, int s_len, int d_len, vec_p scratch) {                     \
        (void)scratch;                                                                                    \
        gather<T>(d, s, i, nullptr, d_len, s_len, no_seg());                                              \
    }                                                                                                     \
    int bck_pu##TY##_scratch(int s_len, int d_len) { (void)s_len; (void)d_len; return 0; }                \
    unsigned int bck_pu##TY##_inplace(void) { return INPLACE_2; }                                         \
    void bck_pe##TY(vec_p d, vec_p s, vec_p i, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d,  \
                    vec_p scratch) {                                                                      \
        (void)scratch;                                                                                    \
        if (m_s <= 1) { gather<T>(d, s, i, nullptr, n_d, n_s, no_seg()); return; }                         \
        SegdView src(sd_s, n_s, m_s);                                                                     \
        gather<T>(d, s, i, nullptr, n_d, n_s, SegBase{SegdView(sd_d, n_d, m_d), src.offs, true});         \
    }                                                                                                     \
    int bck_pe##TY##_scratch(int a, int b, int c, int e) { (void)a; (void)b; (void)c; (void)e; return 0; } \
    unsigned int bck_pe##TY##_inplace(void) { return INPLACE_2; }                                         \
    void fpm_pu##TY(vec_p d, vec_p s, vec_p i, vec_p f, int len_src, int len_dest, vec_p scratch) {        \
        scatter_lww<T>(d, s, i, f, len_src, len_dest, no_seg(), scratch);                                 \
    }                                                                                                     \
    int fpm_pu##TY##_scratch(int s_len, int d_len) { (void)s_len; return lww_units(d_len); }             \
    unsigned int fpm_pu##TY##_inplace(void) { return INPLACE_NONE; }                                      \
    void fpm_pe##TY(vec_p d, vec_p s, vec_p i, vec_p f, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d,  \
                    int m_d, vec_p scratch) {                                                             \
        if (m_s <= 1) { scatter_lww<T>(d, s, i, f, n_s, n_d, no_seg(), scratch); return; }                 \
        SegdView dst(sd_d, n_d, m_d);                                                                     \
        scatter_lww<T>(d, s, i, f, n_s, n_d, SegBase{SegdView(sd_s, n_s, m_s), dst.offs, true}, scratch);  \
    }                                                                                                     \
    int fpm_pe##TY##_scratch(int a, int b, int c, int e) { (void)a; (void)b; (void)e; return lww_units(c); } \
    unsigned int fpm_pe##TY##_inplace(void) { return INPLACE_NONE; }                                      \
    void bfp_pu##TY(vec_p d, vec_p s, vec_p i, vec_p f, int len_src, int len_dest, vec_p scratch) {        \
        (void)scratch;                                                                                    \
        gather<T>(d, s, i, f, len_dest, len_src, no_seg());                                               \
    }                                                                                                     \
    int bfp_pu##TY##_scratch(int s_len, int d_len) { (void)s_len; (void)d_len; return 0; }                \
    unsigned int bfp_pu##TY##_inplace(void) { return INPLACE_NONE; }                                      \
    void bfp_pe##TY(vec_p d, vec_p s, vec_p i, vec_p f, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d,  \
                    int m_d, vec_p scratch) {                                                             \
        (void)scratch;                                                                                    \
        if (m_s <= 1) { gather<T>(d, s, i, f, n_d, n_s, no_seg()); return; }                               \
        SegdView src(sd_s, n_s, m_s);                                                                     \
        gather<T>(d, s, i, f, n_d, n_s, SegBase{SegdView(sd_d, n_d, m_d), src.offs, true});               \
    }                                                                                                     \
    int bfp_pe##TY##_scratch(int a, int b, int c, int e) { (void)a; (void)b; (void)c; (void)e; return 0; } \
    unsigned int bfp_pe##TY##_inplace(void) { return INPLACE_NONE; }                                      \
    void dpe_pu##TY(vec_p d, vec_p s, vec_p i, vec_p v, int len_src, int len_dest, vec_p scratch) {        \
        copy_default(d, v, len_dest, W);                                                                  \
        scatter_lww<T>(d, s, i, nullptr, len_src, len_dest, no_seg(), scratch);                           \
    }                                                                                                     \
    int dpe_pu##TY##_scratch(int s_len, int d_len) { (void)s_len; return lww_units(d_len); }             \
    unsigned int dpe_pu##TY##_inplace(void) { return INPLACE_3; }                                         \
    void dpe_pe##TY(vec_p d, vec_p s, vec_p i, vec_p v, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d,  \
                    int m_d, vec_p scratch) {                                                             \
        copy_default(d, v, n_d, W);                                                                       \
        if (m_s <= 1) { scatter_lww<T>(d, s, i, nullptr, n_s, n_d, no_seg(), scratch); return; }           \
        SegdView dst(sd_d, n_d, m_d);                                                                     \
        scatter_lww<T>(d, s, i, nullptr, n_s, n_d, SegBase{SegdView(sd_s, n_s, m_s), dst.offs, true}, scratch); \
    }                                                                                                     \
    int dpe_pe##TY##_scratch(int a, int b, int c, int e) { (void)a; (void)b; (void)e; return lww_units(c); } \
    unsigned int dpe_pe##TY##_inplace(void) { return INPLACE_3; }                                         \
    void dfp_pu##TY(vec_p d, vec_p s, vec_p i, vec_p f, vec_p v, int len_src, int len_dest, vec_p scratch) { \
        copy_default(d, v, len_dest, W);                                                                  \
        scatter_lww<T>(d, s, i, f, len_src, len_dest, no_seg(), scratch);                                 \
    }                                                                                                     \
    int dfp_pu##TY##_scratch(int s_len, int d_len) { (void)s_len; return lww_units(d_len); }             \
    unsigned int dfp_pu##TY##_inplace(void) { return INPLACE_4; }                                         \
    void dfp_pe##TY(vec_p d, vec_p s, vec_p i, vec_p f, vec_p v, vec_p sd_s, int n_s, int m_s, vec_p sd_d,  \
                    int n_d, int m_d, vec_p scratch) {                                                    \
        copy_default(d, v, n_d, W);                                                                       \
        if (m_s <= 1) { scatter_lww<T>(d, s, i, f, n_s, n_d, no_seg(), scratch); return; }                 \
        SegdView dst(sd_d, n_d, m_d);                                                                     \
        scatter_lww<T>(d, s, i, f, n_s, n_d, SegBase{SegdView(sd_s, n_s, m_s), dst.offs, true}, scratch);  \
    }                                                                                                     \
    int dfp_pe##TY##_scratch(int a, int b, int c, int e) { (void)a; (void)b; (void)e; return lww_units(c); } \
    unsigned int dfp_pe##TY##_inplace(void) { return INPLACE_4; }

extern "C" {
PERMS(z, int, 4)
PERMS(b, int, 4)
PERMS(d, double, 8)
}
