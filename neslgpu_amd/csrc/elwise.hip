// elwise.hip -- elementwise operations, distribute, index, extract,
// replace, copies and RAND (ref nesl/cvl/serial/elwise.c:128-298,
// vprims.c:203-344, library.c:9-49, facilt.c:292-312).
//
// Elementwise kernels are grid-stride loops (coalesced 4- or 8-byte lane
// accesses); one read of each input and one write of the output.  Arithmetic follows C on the
// host: int overflow wraps, int / and % truncate, floating point is compiled
// without contraction (-ffp-contract=off) so + - * / and sqrt round exactly
// as the serial CVL; libm transcendentals may differ in the last ulp.
#include "cvl_hip.hpp"
#include "elwise_ops.hpp"
#include "segd.hpp"
#include "../../include/cvl.h"
#include "../../include/cvl_hip_ext.h"

#include <math.h>

namespace cvlhip {


static int ew_grid(int64_t n) {
    int g = grid_for(n, kBlock * 4);
    return g > 16384 ? 16384 : g;
}

// ------------------------------------------------------------------ kernels
template <class F>
__global__ void k_ew2(typename F::Out *__restrict__ d, const typename F::In *__restrict__ a,
                      const typename F::In *__restrict__ b, int64_t n) {
    const F f;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = f(a[i], b[i]);
}

template <class F>
__global__ void k_ew1(typename F::Out *__restrict__ d, const typename F::In *__restrict__ a, int64_t n) {
    const F f;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = f(a[i]);
}

template <class T>
__global__ void k_sel(T *__restrict__ d, const int *__restrict__ f, const T *__restrict__ a,
                      const T *__restrict__ b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = f[i] ? a[i] : b[i];
}

// one element written by one lane (scalar replace without a host wait)
template <class T>
__global__ void k_put1(T *p, T v) {
    if (threadIdx.x == 0) *p = v;
}

// One element to the host mailbox (ext_vu*).
template <class T>
__global__ void k_ext_post(const T *p, Mailbox *mb, unsigned long long seq) {
    if (threadIdx.x == 0) mailbox_post(mb, seq, scalar_bits(*p));
}

template <class T>
__global__ void k_dist(T *__restrict__ d, T v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = v;
}

template <class T>
__global__ void k_dist_seg(T *__restrict__ d, const T *__restrict__ v, SegdView sv) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (i >= sv.n) return;
    const bool ne = sv.all_nonempty();
    // four elements per step, their lookups in flight together: each is a
    // chain of two dependent loads (bitmap word, then the segment's value)
    for (; i < sv.n; i += 4 * stride) {
        int g[4];
        T x[4];
#pragma unroll
        for (int u = 0; u < 4; u++) g[u] = i + u * stride < sv.n ? sv.seg_of(i + u * stride, ne) : 0;
#pragma unroll
        for (int u = 0; u < 4; u++) x[u] = i + u * stride < sv.n ? v[g[u]] : T(0);
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (i + u * stride < sv.n) d[i + u * stride] = x[u];
    }
}

__global__ void k_index(int *__restrict__ d, int init, int stride_v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = (int)((unsigned)init + (unsigned)stride_v * (unsigned)i);
}

__global__ void k_index_seg(int *__restrict__ d, const int *__restrict__ init, const int *__restrict__ st,
                            SegdView sv) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (i >= sv.n) return;
    const bool ne = sv.all_nonempty();
    for (; i < sv.n; i += 4 * stride) {        // four elements per step, as in k_dist_seg
        int g[4];
#pragma unroll
        for (int u = 0; u < 4; u++) g[u] = i + u * stride < sv.n ? sv.seg_of(i + u * stride, ne) : 0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t e = i + u * stride;
            if (e < sv.n) {
                const unsigned k = (unsigned)(e - sv.offs[g[u]]);
                d[e] = (int)((unsigned)init[g[u]] + (unsigned)st[g[u]] * k);
            }
        }
    }
}

template <class T>
__global__ void k_ext_seg(T *__restrict__ d, const T *__restrict__ s, const int *__restrict__ idx, SegdView sv,
                          unsigned *err) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; g < sv.m; g += (int64_t)gridDim.x * blockDim.x) {
        int64_t p = (int64_t)sv.offs[g] + idx[g];
        if (p >= 0 && p < sv.n) d[g] = s[p];
        else *err = 1u;
    }
}

template <class T>
__global__ void k_rep_seg(T *__restrict__ d, const int *__restrict__ idx, const T *__restrict__ v, SegdView sv,
                          unsigned *err) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; g < sv.m; g += (int64_t)gridDim.x * blockDim.x) {
        int64_t p = (int64_t)sv.offs[g] + idx[g];
        if (p >= 0 && p < sv.n) d[p] = v[g];
        else *err = 1u;
    }
}

// ------------------------------------------------------------------ RAND
// lrand48: X' = (a X + c) mod 2^48, value = X' >> 17 (POSIX).  Element k of
// a call uses the (k+1)-th state after the call's starting state; each
// thread jumps to its first state with the affine map raised to its offset.
static uint64_t g_rand_state = 0;   // glibc's unseeded lrand48 starts at 0
constexpr uint64_t kLcgA = 0x5DEECE66DULL, kLcgC = 0xBULL, kMask48 = (1ULL << 48) - 1;

struct Affine { uint64_t a, c; };   // x -> a x + c (mod 2^48)

__host__ __device__ inline Affine compose(Affine f, Affine g) {   // g after f
    return Affine{(g.a * f.a) & kMask48, (g.a * f.c + g.c) & kMask48};
}
__host__ __device__ inline Affine lcg_pow(uint64_t k) {
    Affine r{1, 0}, b{kLcgA, kLcgC};
    while (k) {
        if (k & 1) r = compose(r, b);
        b = compose(b, b);
        k >>= 1;
    }
    return r;
}

constexpr int kRandPer = 16;

__global__ void k_rand(int *__restrict__ d, const int *__restrict__ s, int64_t n, uint64_t x0) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t k0 = t * kRandPer;
    if (k0 >= n) return;
    Affine j = lcg_pow((uint64_t)k0);
    uint64_t x = (j.a * x0 + j.c) & kMask48;
    for (int u = 0; u < kRandPer && k0 + u < n; u++) {
        x = (kLcgA * x + kLcgC) & kMask48;
        int r = (int)(x >> 17);
        int m = s[k0 + u];
        d[k0 + u] = m != 0 ? r % m : 0;
    }
}

// ------------------------------------------------------------------ splitmix
__device__ __forceinline__ uint64_t sm64(uint64_t seed, uint64_t k) {
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ void k_splitmix_u32(uint32_t *d, int64_t n, uint64_t seed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = (uint32_t)(sm64(seed, (uint64_t)i) >> 32);
}

__global__ void k_splitmix_f64(double *d, int64_t n, uint64_t seed, double lo, double hi) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double u = (double)(sm64(seed, (uint64_t)i) >> 11) * 0x1.0p-53;
        d[i] = lo + (hi - lo) * u;
    }
}

}  // namespace cvlhip

using namespace cvlhip;

// ------------------------------------------------------------------ CVL entry points
#define NOSCR(NAME, FLAG)                                     \
    int NAME##_scratch(int len) { (void)len; return 0; }      \
    unsigned int NAME##_inplace(void) { return FLAG; }

#define EW2(NAME, F, FLAG)                                                                    \
    void NAME(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch) {                          \
        (void)scratch;                                                                        \
        if (len <= 0) return;                                                                 \
        k_ew2<F><<<ew_grid(len), kBlock, 0, g_stream>>>((F::Out *)d, (const F::In *)s1,       \
                                                        (const F::In *)s2, len);              \
        LAUNCHCHECK();                                                                        \
    }                                                                                         \
    NOSCR(NAME, FLAG)

#define EW1(NAME, F, FLAG)                                                                    \
    void NAME(vec_p d, vec_p s, int len, vec_p scratch) {                                     \
        (void)scratch;                                                                        \
        if (len <= 0) return;                                                                 \
        k_ew1<F><<<ew_grid(len), kBlock, 0, g_stream>>>((F::Out *)d, (const F::In *)s, len);  \
        LAUNCHCHECK();                                                                        \
    }                                                                                         \
    NOSCR(NAME, FLAG)

// In-place flags as in the serial CVL (elwise.c:128-177): an output may
// overwrite any input of the same width.
#define IP12 (INPLACE_1 | INPLACE_2)

extern "C" {

EW2(max_wuz, MaxZf, IP12) EW2(max_wud, MaxDf, IP12)
EW2(min_wuz, MinZf, IP12) EW2(min_wud, MinDf, IP12)
EW2(add_wuz, AddZf, IP12) EW2(add_wud, AddDf, IP12)
EW2(sub_wuz, SubZf, IP12) EW2(sub_wud, SubDf, IP12)
EW2(mul_wuz, MulZf, IP12) EW2(mul_wud, MulDf, IP12)
EW2(div_wuz, DivZf, IP12) EW2(div_wud, DivDf, IP12)
EW2(grt_wuz, GrtZf, IP12) EW2(grt_wud, GrtDf, INPLACE_NONE)
EW2(les_wuz, LesZf, IP12) EW2(les_wud, LesDf, INPLACE_NONE)
EW2(geq_wuz, GeqZf, IP12) EW2(geq_wud, GeqDf, INPLACE_NONE)
EW2(leq_wuz, LeqZf, IP12) EW2(leq_wud, LeqDf, INPLACE_NONE)
EW2(lsh_wuz, LshZf, IP12)
EW2(rsh_wuz, RshZf, IP12)
EW2(mod_wuz, ModZf, IP12)
EW2(eql_wub, EqlBf, IP12) EW2(eql_wuz, EqlZf, IP12) EW2(eql_wud, EqlDf, INPLACE_NONE)
EW2(neq_wub, NeqBf, IP12) EW2(neq_wuz, NeqZf, IP12) EW2(neq_wud, NeqDf, INPLACE_NONE)
EW2(xor_wub, XorBf, IP12)
EW2(ior_wub, IorBf, IP12)
EW2(and_wub, AndBf, IP12)
EW2(ior_wuz, IorZf, IP12)
EW2(and_wuz, AndZf, IP12)
EW2(xor_wuz, XorZf, IP12)

EW1(not_wub, NotBf, INPLACE_1)
EW1(not_wuz, NotZf, INPLACE_1)
EW1(int_wud, IntDf, INPLACE_NONE)
EW1(int_wub, IntBf, INPLACE_1)
EW1(dbl_wuz, DblZf, INPLACE_NONE)
EW1(boo_wuz, BooZf, INPLACE_1)
EW1(flr_wud, FlrDf, INPLACE_NONE)
EW1(cei_wud, CeiDf, INPLACE_NONE)
EW1(trn_wud, TrnDf, INPLACE_NONE)
EW1(rou_wud, RouDf, INPLACE_NONE)
EW1(exp_wud, ExpDf, INPLACE_1)
EW1(log_wud, LogDf, INPLACE_1)
EW1(sqt_wud, SqtDf, INPLACE_1)
EW1(sin_wud, SinDf, INPLACE_1)
EW1(cos_wud, CosDf, INPLACE_1)
EW1(tan_wud, TanDf, INPLACE_1)
EW1(asn_wud, AsnDf, INPLACE_1)
EW1(acs_wud, AcsDf, INPLACE_1)
EW1(atn_wud, AtnDf, INPLACE_1)
EW1(snh_wud, SnhDf, INPLACE_1)
EW1(csh_wud, CshDf, INPLACE_1)
EW1(tnh_wud, TnhDf, INPLACE_1)

#define CPY(NAME, W)                                                                           \
    void NAME(vec_p d, vec_p s, int len, vec_p scratch) {                                      \
        (void)scratch;                                                                         \
        if (len <= 0 || d == s) return;                                                        \
        HIPCHECK(hipMemcpyAsync(d, s, (size_t)len * W, hipMemcpyDeviceToDevice, g_stream));    \
    }                                                                                          \
    NOSCR(NAME, INPLACE_1)
CPY(cpy_wuz, 4)
CPY(cpy_wub, 4)
CPY(cpy_wud, 8)

#define SEL(NAME, T)                                                                           \
    void NAME(vec_p d, vec_p s1, vec_p s2, vec_p s3, int len, vec_p scratch) {                 \
        (void)scratch;                                                                         \
        if (len <= 0) return;                                                                  \
        k_sel<T><<<ew_grid(len), kBlock, 0, g_stream>>>((T *)d, (const int *)s1, (const T *)s2, (const T *)s3, len); \
        LAUNCHCHECK();                                                                         \
    }                                                                                          \
    NOSCR(NAME, INPLACE_1 | INPLACE_2 | INPLACE_3)
SEL(sel_wuz, int)
SEL(sel_wub, int)
SEL(sel_wud, double)

// ---- RAND
void rnd_wuz(vec_p d, vec_p s, int len, vec_p scratch) {
    (void)scratch;
    if (len <= 0) return;
    int64_t threads = ((int64_t)len + kRandPer - 1) / kRandPer;
    k_rand<<<grid_for(threads, kBlock), kBlock, 0, g_stream>>>((int *)d, (const int *)s, len, g_rand_state);
    LAUNCHCHECK();
    Affine j = lcg_pow((uint64_t)len);
    g_rand_state = (j.a * g_rand_state + j.c) & kMask48;
}
NOSCR(rnd_wuz, INPLACE_1)

// rnd_foz: the serial CVL on Linux calls srandom(seed) (facilt.c:292-312),
// which does not reseed lrand48; RAND's stream is therefore unaffected.
void rnd_foz(int seed) { (void)seed; }

void cvlhip_rand_set_state(unsigned long long s) { g_rand_state = s & kMask48; }
unsigned long long cvlhip_rand_get_state(void) { return g_rand_state; }

// ---- distribute / index
#define DIS(NAME, T)                                                                           \
    void NAME(vec_p d, T v, int len, vec_p scratch) {                                          \
        (void)scratch;                                                                         \
        if (len <= 0) return;                                                                  \
        k_dist<T><<<ew_grid(len), kBlock, 0, g_stream>>>((T *)d, v, len);                      \
        LAUNCHCHECK();                                                                         \
    }                                                                                          \
    NOSCR(NAME, INPLACE_NONE)
DIS(dis_vuz, int)
DIS(dis_vud, double)
void dis_vub(vec_p d, cvl_bool v, int len, vec_p scratch) {
    (void)scratch;
    if (len <= 0) return;
    k_dist<int><<<ew_grid(len), kBlock, 0, g_stream>>>((int *)d, v, len);
    LAUNCHCHECK();
}
NOSCR(dis_vub, INPLACE_NONE)

#define DISSEG(NAME, T, UNSEG)                                                                 \
    void NAME(vec_p d, vec_p v, vec_p sd, int n, int m, vec_p scratch) {                       \
        (void)scratch;                                                                         \
        if (n <= 0) return;                                                                    \
        SegdView sv(sd, n, m);                                                                 \
        k_dist_seg<T><<<ew_grid(n), kBlock, 0, g_stream>>>((T *)d, (const T *)v, sv);          \
        LAUNCHCHECK();                                                                         \
    }                                                                                          \
    int NAME##_scratch(int n, int m) { (void)n; (void)m; return 0; }                           \
    unsigned int NAME##_inplace(void) { return INPLACE_NONE; }
DISSEG(dis_vez, int, dis_vuz)
DISSEG(dis_veb, int, dis_vub)
DISSEG(dis_ved, double, dis_vud)

void ind_luz(vec_p d, int init, int stride, int count, vec_p scratch) {
    (void)scratch;
    if (count <= 0) return;
    k_index<<<ew_grid(count), kBlock, 0, g_stream>>>((int *)d, init, stride, count);
    LAUNCHCHECK();
}
NOSCR(ind_luz, INPLACE_NONE)

void ind_lez(vec_p d, vec_p init, vec_p stride, vec_p dest_segd, int vec_len, int seg_count, vec_p scratch) {
    (void)scratch;
    if (vec_len <= 0) return;
    SegdView sv(dest_segd, vec_len, seg_count);
    k_index_seg<<<ew_grid(vec_len), kBlock, 0, g_stream>>>((int *)d, (const int *)init, (const int *)stride, sv);
    LAUNCHCHECK();
}
int ind_lez_scratch(int n, int m) { (void)n; (void)m; return 0; }
unsigned int ind_lez_inplace(void) { return INPLACE_NONE; }

// ---- extract / replace of one element
// EXT has a one-lane kernel post the element to the host mailbox and waits
// for it (the caller needs the value); REP writes it with a one-lane kernel
// that takes the value as an argument, so it queues behind earlier work
// without a host wait.
#define EXT(NAME, T)                                                                           \
    T NAME(vec_p V, int i, int len, vec_p scratch) {                                           \
        (void)scratch;                                                                         \
        if (i < 0 || i >= len) { usage_error(#NAME ": index %d outside [0,%d)", i, len); return (T)0; } \
        Mailbox *mb = mailbox_device();                                                        \
        const unsigned long long seq = mailbox_next();                                         \
        k_ext_post<T><<<1, 64, 0, g_stream>>>((const T *)V + i, mb, seq);                      \
        LAUNCHCHECK();                                                                         \
        return scalar_from_bits<T>(mailbox_wait(seq));                                         \
    }                                                                                          \
    NOSCR(NAME, INPLACE_NONE)
EXT(ext_vuz, int)
EXT(ext_vub, cvl_bool)
EXT(ext_vud, double)

#define REP(NAME, T, NORM)                                                                     \
    void NAME(vec_p V, int i, T val, int len, vec_p scratch) {                                 \
        (void)scratch;                                                                         \
        if (i < 0 || i >= len) { usage_error(#NAME ": index %d outside [0,%d)", i, len); return; } \
        k_put1<T><<<1, 64, 0, g_stream>>>((T *)V + i, (T)NORM(val));                          \
        LAUNCHCHECK();                                                                         \
    }                                                                                          \
    NOSCR(NAME, INPLACE_NONE)
#define IDENT(x) (x)
#define NOTNOT(x) ((x) != 0)
REP(rep_vuz, int, IDENT)
REP(rep_vub, cvl_bool, NOTNOT)
REP(rep_vud, double, IDENT)

#define EXTSEG(NAME, T)                                                                        \
    void NAME(vec_p d, vec_p s, vec_p i, vec_p sd, int n, int m, vec_p scratch) {              \
        (void)scratch;                                                                         \
        if (m <= 0) return;                                                                    \
        SegdView sv(sd, n, m);                                                                 \
        k_ext_seg<T><<<grid_for(m, kBlock), kBlock, 0, g_stream>>>((T *)d, (const T *)s, (const int *)i, sv, \
                                                                   device_error_word());       \
        LAUNCHCHECK();                                                                         \
    }                                                                                          \
    int NAME##_scratch(int n, int m) { (void)n; (void)m; return 0; }                           \
    unsigned int NAME##_inplace(void) { return INPLACE_2; }
EXTSEG(ext_vez, int)
EXTSEG(ext_veb, int)
EXTSEG(ext_ved, double)

#define REPSEG(NAME, T)                                                                        \
    void NAME(vec_p d, vec_p s, vec_p v, vec_p sd, int n, int m, vec_p scratch) {              \
        (void)scratch;                                                                         \
        if (m <= 0) return;                                                                    \
        SegdView sv(sd, n, m);                                                                 \
        k_rep_seg<T><<<grid_for(m, kBlock), kBlock, 0, g_stream>>>((T *)d, (const int *)s, (const T *)v, sv, \
                                                                   device_error_word());       \
        LAUNCHCHECK();                                                                         \
    }                                                                                          \
    int NAME##_scratch(int n, int m) { (void)n; (void)m; return 0; }                           \
    unsigned int NAME##_inplace(void) { return INPLACE_NONE; }
REPSEG(rep_vez, int)
REPSEG(rep_veb, int)
REPSEG(rep_ved, double)

// ---- synthetic inputs
void cvlhip_fill_splitmix_u32(vec_p d, long long n, unsigned long long seed) {
    if (n <= 0) return;
    k_splitmix_u32<<<ew_grid(n), kBlock, 0, g_stream>>>((uint32_t *)d, n, seed);
    LAUNCHCHECK();
}
void cvlhip_fill_splitmix_f64(vec_p d, long long n, unsigned long long seed, double lo, double hi) {
    if (n <= 0) return;
    k_splitmix_f64<<<ew_grid(n), kBlock, 0, g_stream>>>((double *)d, n, seed, lo, hi);
    LAUNCHCHECK();
}

}  // extern "C"
