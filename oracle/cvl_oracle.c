/*
 * cvl_oracle.c -- CPU restatement of the serial CVL semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * backend in neslgpu_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * Every function restates the algorithm of the reference serial CVL
 * (nesl/cvl/serial/*.c) as plain sequential C.  Each block cites the
 * reference lines it follows.  Build: gcc -O2 -fwrapv (see oracle/Makefile);
 * integer arithmetic that may overflow is done on uint32_t so wraparound is
 * defined whatever the flags.
 *
 * How each part is pinned (tests/test_oracle.py):
 *   - scans/reductions (int, bool, double): golden vectors from the compiled
 *     reference serial CVL (tests/golden), brute-force closed forms on tiny
 *     inputs, the test.nesl known-answer tests;
 *   - pack / permute family / index / dist / ext / rep: golden vectors plus
 *     invariants (pack count == reduce(flags), permute o inverse == id);
 *   - rank (int, double, up/down, segmented): golden vectors, the test.nesl
 *     rank KATs, and a stable numpy argsort on tiny inputs;
 *   - lrand48: glibc's own lrand48()/srand48() in a fresh process;
 *   - Neumaier scan: exact rational arithmetic (Python fractions) on small
 *     inputs.
 */
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <math.h>

/* ------------------------------------------------------------------ */
/* Types, operators and identities                                     */
/* ------------------------------------------------------------------ */

/* Operator codes (shared with the HIP backend's Python mirror). */
enum { ORC_ADD = 0, ORC_MUL = 1, ORC_MIN = 2, ORC_MAX = 3,
       ORC_AND = 4, ORC_IOR = 5, ORC_XOR = 6 };
/* Element types: z = int32, d = double, b = cvl_bool (int32 0/1). */
enum { ORC_Z = 0, ORC_D = 1, ORC_B = 2 };

/* Identities: ref nesl/cvl/serial/defins.h:218-228 (MAX_INT, MIN_INT and
 * the float-rounded +-1e36 used for doubles), vprims.c:30-49 (per op). */
#define ORC_MAX_INT ((int32_t)0x7fffffff)
#define ORC_MIN_INT ((int32_t)0x80000000)
#define ORC_MAX_DBL ((double)(float)1.0e36)
#define ORC_MIN_DBL ((double)(float)-1.0e36)

static inline int32_t z_op(int op, int32_t a, int32_t b)
{
    switch (op) {
    case ORC_ADD: return (int32_t)((uint32_t)a + (uint32_t)b);
    case ORC_MUL: return (int32_t)((uint32_t)a * (uint32_t)b);
    case ORC_MIN: return a < b ? a : b;            /* defins.h:62 */
    case ORC_MAX: return a > b ? a : b;            /* defins.h:61 */
    case ORC_AND: return a & b;                    /* band */
    case ORC_IOR: return a | b;                    /* bor  */
    default:      return a ^ b;                    /* xor  */
    }
}
static inline int32_t z_id(int op)
{
    switch (op) {
    case ORC_ADD: return 0;
    case ORC_MUL: return 1;
    case ORC_MIN: return ORC_MAX_INT;
    case ORC_MAX: return ORC_MIN_INT;
    case ORC_AND: return ~0;
    default:      return 0;
    }
}
static inline double d_op(int op, double a, double b)
{
    switch (op) {
    case ORC_ADD: return a + b;
    case ORC_MUL: return a * b;
    case ORC_MIN: return a < b ? a : b;
    default:      return a > b ? a : b;
    }
}
static inline double d_id(int op)
{
    switch (op) {
    case ORC_ADD: return 0.0;
    case ORC_MUL: return 1.0;
    case ORC_MIN: return ORC_MAX_DBL;
    default:      return ORC_MIN_DBL;
    }
}
/* Logical ops on cvl_bool: && and || give 0/1; lxor is (!!a)^(!!b)
 * (defins.h:99-106). */
static inline int32_t b_op(int op, int32_t a, int32_t b)
{
    switch (op) {
    case ORC_AND: return (a && b) ? 1 : 0;
    case ORC_IOR: return (a || b) ? 1 : 0;
    default:      return (!!a) ^ (!!b);
    }
}
static inline int32_t b_id(int op) { return op == ORC_AND ? 1 : 0; }

static int op_ok(int op, int ty)
{
    if (ty == ORC_Z) return op >= ORC_ADD && op <= ORC_XOR;
    if (ty == ORC_D) return op >= ORC_ADD && op <= ORC_MAX;
    if (ty == ORC_B) return op >= ORC_AND && op <= ORC_XOR;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Scans and reductions                                                */
/* ------------------------------------------------------------------ */

/* Exclusive scan of one run: out[k] = op(id, x[0..k-1]) in left-to-right
 * order.  Unsegmented scans: ref vprims.c:17-49 (simpscan).  The whole
 * vector is one run. */
#define ORC_RUN(T, OPF, ID)                                              \
    do {                                                                 \
        T acc_ = (ID);                                                   \
        for (int64_t k_ = 0; k_ < len_; k_++) {                          \
            T x_ = src_[k_];                                             \
            dst_[k_] = acc_;                                             \
            acc_ = OPF(op, acc_, x_);                                    \
        }                                                                \
    } while (0)

int orc_scan(int op, int ty, void *d, const void *s, int64_t n)
{
    if (!op_ok(op, ty)) return -1;
    int64_t len_ = n;
    if (ty == ORC_D) {
        double *dst_ = (double *)d; const double *src_ = (const double *)s;
        ORC_RUN(double, d_op, d_id(op));
    } else if (ty == ORC_Z) {
        int32_t *dst_ = (int32_t *)d; const int32_t *src_ = (const int32_t *)s;
        ORC_RUN(int32_t, z_op, z_id(op));
    } else {
        int32_t *dst_ = (int32_t *)d; const int32_t *src_ = (const int32_t *)s;
        ORC_RUN(int32_t, b_op, b_id(op));
    }
    return 0;
}

/* Segmented exclusive scan: every segment restarts from the identity.
 * ref vprims.c:62-109 (simpsegscan); the m == 1 shortcut to the
 * unsegmented scan (vprims.c:75) gives the same result. */
int orc_segscan(int op, int ty, void *d, const void *s, const int32_t *lens,
                int64_t m)
{
    if (!op_ok(op, ty)) return -1;
    int64_t base = 0;
    for (int64_t g = 0; g < m; g++) {
        int64_t len_ = lens[g];
        if (ty == ORC_D) {
            double *dst_ = (double *)d + base;
            const double *src_ = (const double *)s + base;
            ORC_RUN(double, d_op, d_id(op));
        } else if (ty == ORC_Z) {
            int32_t *dst_ = (int32_t *)d + base;
            const int32_t *src_ = (const int32_t *)s + base;
            ORC_RUN(int32_t, z_op, z_id(op));
        } else {
            int32_t *dst_ = (int32_t *)d + base;
            const int32_t *src_ = (const int32_t *)s + base;
            ORC_RUN(int32_t, b_op, b_id(op));
        }
        base += len_;
    }
    return 0;
}

/* Reduction: fold from the identity, left to right.  ref vprims.c:114-146.
 * The result is written to *out (int32 or double). */
int orc_reduce(int op, int ty, void *out, const void *s, int64_t n)
{
    if (!op_ok(op, ty)) return -1;
    if (ty == ORC_D) {
        const double *x = (const double *)s; double a = d_id(op);
        for (int64_t k = 0; k < n; k++) a = d_op(op, a, x[k]);
        *(double *)out = a;
    } else if (ty == ORC_Z) {
        const int32_t *x = (const int32_t *)s; int32_t a = z_id(op);
        for (int64_t k = 0; k < n; k++) a = z_op(op, a, x[k]);
        *(int32_t *)out = a;
    } else {
        const int32_t *x = (const int32_t *)s; int32_t a = b_id(op);
        for (int64_t k = 0; k < n; k++) a = b_op(op, a, x[k]);
        *(int32_t *)out = a;
    }
    return 0;
}

/* Segmented reduction: one value per segment, the identity for an empty
 * segment.  ref vprims.c:155-201 (segreduce). */
int orc_segreduce(int op, int ty, void *d, const void *s, const int32_t *lens,
                  int64_t m)
{
    if (!op_ok(op, ty)) return -1;
    int64_t base = 0;
    size_t w = ty == ORC_D ? 8 : 4;
    for (int64_t g = 0; g < m; g++) {
        orc_reduce(op, ty, (char *)d + g * w, (const char *)s + base * w,
                   lens[g]);
        base += lens[g];
    }
    return 0;
}

/* Compensated (Neumaier) exclusive segmented +-scan of doubles.  Not a
 * reference function: it is the accuracy yardstick for float +-scans,
 * whose summation order the GPU changes.  out[k] is the compensated
 * prefix (sum + compensation) and absp[k] the running sum of |x| of the
 * same prefix, from which the tolerance is formed. */
void orc_segscan_add_neumaier(double *out, double *absp, const double *s,
                              const int32_t *lens, int64_t m)
{
    int64_t base = 0;
    for (int64_t g = 0; g < m; g++) {
        double sum = 0.0, c = 0.0, a = 0.0;
        for (int64_t k = 0; k < lens[g]; k++) {
            double x = s[base + k];
            out[base + k] = sum + c;
            absp[base + k] = a;
            double t = sum + x;
            if (fabs(sum) >= fabs(x)) c += (sum - t) + x;
            else                      c += (x - t) + sum;
            sum = t;
            a += fabs(x);
        }
        base += lens[g];
    }
}

/* Compensated segmented +-reduce (same yardstick, one value per segment). */
void orc_segreduce_add_neumaier(double *out, double *absum, const double *s,
                                const int32_t *lens, int64_t m)
{
    int64_t base = 0;
    for (int64_t g = 0; g < m; g++) {
        double sum = 0.0, c = 0.0, a = 0.0;
        for (int64_t k = 0; k < lens[g]; k++) {
            double x = s[base + k];
            double t = sum + x;
            if (fabs(sum) >= fabs(x)) c += (sum - t) + x;
            else                      c += (x - t) + sum;
            sum = t;
            a += fabs(x);
        }
        out[g] = sum + c;
        absum[g] = a;
        base += lens[g];
    }
}

/* ------------------------------------------------------------------ */
/* Segment descriptors                                                 */
/* ------------------------------------------------------------------ */

/* The serial segment descriptor is the vector of m lengths
 * (ref facilt.c:106-144).  offsets[g] = sum of lens[0..g-1], offsets[m] = n. */
void orc_offsets(int64_t *offs, const int32_t *lens, int64_t m)
{
    int64_t a = 0;
    for (int64_t g = 0; g < m; g++) { offs[g] = a; a += lens[g]; }
    offs[m] = a;
}

/* ------------------------------------------------------------------ */
/* index, distribute, extract, replace                                 */
/* ------------------------------------------------------------------ */

/* ref library.c:12-23: dest[i] = init + i*stride, with int wraparound. */
void orc_ind_luz(int32_t *d, int32_t init, int32_t stride, int64_t n)
{
    uint32_t cur = (uint32_t)init;
    for (int64_t k = 0; k < n; k++) { d[k] = (int32_t)cur; cur += (uint32_t)stride; }
}

/* ref library.c:30-47: one index run per segment. */
void orc_ind_lez(int32_t *d, const int32_t *init, const int32_t *stride,
                 const int32_t *lens, int64_t m)
{
    int64_t base = 0;
    for (int64_t g = 0; g < m; g++) {
        orc_ind_luz(d + base, init[g], stride[g], lens[g]);
        base += lens[g];
    }
}

/* ref vprims.c:318-344: every element of segment g receives v[g]. */
void orc_dis_ve(int ty, void *d, const void *v, const int32_t *lens, int64_t m)
{
    size_t w = ty == ORC_D ? 8 : 4;
    int64_t base = 0;
    for (int64_t g = 0; g < m; g++) {
        for (int64_t k = 0; k < lens[g]; k++)
            memcpy((char *)d + (base + k) * w, (const char *)v + g * w, w);
        base += lens[g];
    }
}

/* ref vprims.c:225-246: d[g] = s[offset_g + i[g]]. */
void orc_ext_ve(int ty, void *d, const void *s, const int32_t *idx,
                const int32_t *lens, int64_t m)
{
    size_t w = ty == ORC_D ? 8 : 4;
    int64_t base = 0;
    for (int64_t g = 0; g < m; g++) {
        memcpy((char *)d + g * w, (const char *)s + (base + idx[g]) * w, w);
        base += lens[g];
    }
}

/* ref vprims.c:271-292: d[offset_g + i[g]] = v[g]. */
void orc_rep_ve(int ty, void *d, const int32_t *idx, const void *v,
                const int32_t *lens, int64_t m)
{
    size_t w = ty == ORC_D ? 8 : 4;
    int64_t base = 0;
    for (int64_t g = 0; g < m; g++) {
        memcpy((char *)d + (base + idx[g]) * w, (const char *)v + g * w, w);
        base += lens[g];
    }
}

/* ------------------------------------------------------------------ */
/* pack                                                                */
/* ------------------------------------------------------------------ */

/* ref library.c:66-75: number of true flags. */
int64_t orc_pk1_luv(const int32_t *f, int64_t n)
{
    int64_t c = 0;
    for (int64_t k = 0; k < n; k++) c += f[k] != 0;
    return c;
}

/* ref library.c:79-97 (unsegmented) and 139-164 (segmented, whose output
 * order is the flat order): stable compaction of flagged elements.
 * Returns the number written. */
int64_t orc_pk2(int ty, void *d, const void *s, const int32_t *f, int64_t n)
{
    size_t w = ty == ORC_D ? 8 : 4;
    int64_t o = 0;
    for (int64_t k = 0; k < n; k++)
        if (f[k]) { memcpy((char *)d + o * w, (const char *)s + k * w, w); o++; }
    return o;
}

/* ref library.c:114-133: per-segment count of true flags (the lengths of
 * the packed result's segment descriptor). */
void orc_pk1_lev(int32_t *ds, const int32_t *f, const int32_t *lens, int64_t m)
{
    int64_t base = 0;
    for (int64_t g = 0; g < m; g++) {
        ds[g] = (int32_t)orc_pk1_luv(f + base, lens[g]);
        base += lens[g];
    }
}

/* ------------------------------------------------------------------ */
/* permute family                                                      */
/* ------------------------------------------------------------------ */

#define W_OF(ty) ((ty) == ORC_D ? (size_t)8 : (size_t)4)
#define PUT(d, i, s, k, w) memcpy((char *)(d) + (int64_t)(i) * (w), \
                                  (const char *)(s) + (int64_t)(k) * (w), (w))

/* ref vprims.c:356-371: d[i[k]] = s[k]. */
void orc_smp_pu(int ty, void *d, const void *s, const int32_t *idx, int64_t n)
{
    size_t w = W_OF(ty);
    for (int64_t k = 0; k < n; k++) PUT(d, idx[k], s, k, w);
}

/* ref vprims.c:379-405: index relative to the element's own segment. */
void orc_smp_pe(int ty, void *d, const void *s, const int32_t *idx,
                const int32_t *lens, int64_t m)
{
    size_t w = W_OF(ty);
    int64_t base = 0;
    for (int64_t g = 0; g < m; g++) {
        for (int64_t k = base; k < base + lens[g]; k++)
            PUT(d, base + idx[k], s, k, w);
        base += lens[g];
    }
}

/* ref vprims.c:416-431: d[k] = s[i[k]] for k < n_d. */
void orc_bck_pu(int ty, void *d, const void *s, const int32_t *idx, int64_t n_d)
{
    size_t w = W_OF(ty);
    for (int64_t k = 0; k < n_d; k++) PUT(d, k, s, idx[k], w);
}

/* ref vprims.c:440-467: destination segment g gathers from source
 * segment g: d[k] = s[src_off_g + i[k]].  The loop runs over the m_d
 * destination segments (m_s == 1 takes the unsegmented path, which is the
 * same gather). */
void orc_bck_pe(int ty, void *d, const void *s, const int32_t *idx,
                const int32_t *lens_s, int64_t m_s, const int32_t *lens_d,
                int64_t m_d, int64_t n_d)
{
    size_t w = W_OF(ty);
    if (m_s == 1) { orc_bck_pu(ty, d, s, idx, n_d); return; }
    int64_t sb = 0, db = 0;
    for (int64_t g = 0; g < m_d; g++) {
        for (int64_t k = db; k < db + lens_d[g]; k++) PUT(d, k, s, sb + idx[k], w);
        db += lens_d[g];
        sb += lens_s[g];
    }
}

/* ref vecops.c:22-42: flagged permute, d[i[k]] = s[k] where f[k]. */
void orc_fpm_pu(int ty, void *d, const void *s, const int32_t *idx,
                const int32_t *f, int64_t n_s)
{
    size_t w = W_OF(ty);
    for (int64_t k = 0; k < n_s; k++) if (f[k]) PUT(d, idx[k], s, k, w);
}

/* ref vecops.c:44-73: segmented flagged permute; the destination base of
 * source segment g is the start of destination segment g. */
void orc_fpm_pe(int ty, void *d, const void *s, const int32_t *idx,
                const int32_t *f, const int32_t *lens_s, int64_t m_s,
                const int32_t *lens_d, int64_t m_d, int64_t n_s)
{
    size_t w = W_OF(ty);
    (void)m_d;
    if (m_s == 1) { orc_fpm_pu(ty, d, s, idx, f, n_s); return; }
    int64_t sb = 0, db = 0;
    for (int64_t g = 0; g < m_s; g++) {
        for (int64_t k = sb; k < sb + lens_s[g]; k++)
            if (f[k]) PUT(d, db + idx[k], s, k, w);
        sb += lens_s[g];
        db += lens_d[g];
    }
}

/* ref vecops.c:86-106: flagged back-permute; unflagged positions get 0. */
void orc_bfp_pu(int ty, void *d, const void *s, const int32_t *idx,
                const int32_t *f, int64_t n_d)
{
    size_t w = W_OF(ty);
    for (int64_t k = 0; k < n_d; k++) {
        if (f[k]) PUT(d, k, s, idx[k], w);
        else memset((char *)d + k * w, 0, w);
    }
}

/* ref vecops.c:114-145: per destination segment g (walked with the
 * source segment count), gather from source segment g where flagged. */
void orc_bfp_pe(int ty, void *d, const void *s, const int32_t *idx,
                const int32_t *f, const int32_t *lens_s, int64_t m_s,
                const int32_t *lens_d, int64_t m_d, int64_t n_d)
{
    size_t w = W_OF(ty);
    (void)m_d;
    if (m_s == 1) { orc_bfp_pu(ty, d, s, idx, f, n_d); return; }
    memset(d, 0, (size_t)n_d * w);
    int64_t sb = 0, db = 0;
    for (int64_t g = 0; g < m_s; g++) {
        for (int64_t k = db; k < db + lens_d[g]; k++)
            if (f[k]) PUT(d, k, s, sb + idx[k], w);
        db += lens_d[g];
        sb += lens_s[g];
    }
}

/* ref vecops.c:158-186: copy the default vector, then permute. */
void orc_dpe_pu(int ty, void *d, const void *s, const int32_t *idx,
                const void *v, int64_t n_s, int64_t n_d)
{
    size_t w = W_OF(ty);
    if (d != v) memmove(d, v, (size_t)n_d * w);
    orc_smp_pu(ty, d, s, idx, n_s);
}

/* ref vecops.c:193-219: segmented permute with default. */
void orc_dpe_pe(int ty, void *d, const void *s, const int32_t *idx,
                const void *v, const int32_t *lens_s, int64_t m_s,
                const int32_t *lens_d, int64_t m_d, int64_t n_s, int64_t n_d)
{
    size_t w = W_OF(ty);
    (void)m_d;
    if (m_s == 1) { orc_dpe_pu(ty, d, s, idx, v, n_s, n_d); return; }
    if (d != v) memmove(d, v, (size_t)n_d * w);
    int64_t sb = 0, db = 0;
    for (int64_t g = 0; g < m_s; g++) {
        for (int64_t k = sb; k < sb + lens_s[g]; k++) PUT(d, db + idx[k], s, k, w);
        sb += lens_s[g];
        db += lens_d[g];
    }
}

/* ref vecops.c:241-251 and 262-277: copy default, then flagged permute. */
void orc_dfp_pu(int ty, void *d, const void *s, const int32_t *idx,
                const int32_t *f, const void *v, int64_t n_s, int64_t n_d)
{
    size_t w = W_OF(ty);
    if (d != v) memmove(d, v, (size_t)n_d * w);
    orc_fpm_pu(ty, d, s, idx, f, n_s);
}

void orc_dfp_pe(int ty, void *d, const void *s, const int32_t *idx,
                const int32_t *f, const void *v, const int32_t *lens_s,
                int64_t m_s, const int32_t *lens_d, int64_t m_d, int64_t n_s,
                int64_t n_d)
{
    size_t w = W_OF(ty);
    if (d != v) memmove(d, v, (size_t)n_d * w);
    orc_fpm_pe(ty, d, s, idx, f, lens_s, m_s, lens_d, m_d, n_s);
}

/* ------------------------------------------------------------------ */
/* rank                                                                */
/* ------------------------------------------------------------------ */

/* One stable LSD pass family over 32-bit keys, 8-bit digits, four passes
 * (ref rank.c:49-71, field_rank).  order[] holds a permutation of element
 * indices; on return it lists them stably sorted by key[].  tmp has n
 * entries. */
static void orc_field_rank(int32_t *order, const uint32_t *key, int32_t *tmp,
                           int64_t n)
{
    for (int shift = 0; shift < 32; shift += 8) {
        int64_t cnt[257];
        memset(cnt, 0, sizeof cnt);
        for (int64_t k = 0; k < n; k++) cnt[((key[order[k]] >> shift) & 255) + 1]++;
        for (int b = 0; b < 256; b++) cnt[b + 1] += cnt[b];
        for (int64_t k = 0; k < n; k++) {
            int dg = (key[order[k]] >> shift) & 255;
            tmp[cnt[dg]++] = order[k];
        }
        memcpy(order, tmp, (size_t)n * sizeof(int32_t));
    }
}

/* Final step shared by all ranks (ref rank.c:116-149 and 296-314): an
 * optional stable pass on the segment number, rank[order[p]] = p, then
 * subtract the segment offset so ranks are per segment. */
static void orc_rank_finish(int32_t *rank, int32_t *order, int32_t *tmp,
                            uint32_t *aux, const int32_t *lens, int64_t m,
                            int64_t n)
{
    int seg = lens != NULL && m > 1;
    if (seg) {
        int64_t p = 0;
        for (int64_t g = 0; g < m; g++)
            for (int64_t k = 0; k < lens[g]; k++) aux[p++] = (uint32_t)g;
        orc_field_rank(order, aux, tmp, n);
    }
    for (int64_t p = 0; p < n; p++) rank[order[p]] = (int32_t)p;
    if (seg) {
        int64_t p = 0, base = 0;
        for (int64_t g = 0; g < m; g++) {
            for (int64_t k = 0; k < lens[g]; k++) rank[p++] -= (int32_t)base;
            base += lens[g];
        }
    }
}

/* Integer rank, ref rank.c:88-150 (int_rank_sort) and 152-196 (entry
 * points).  Up: keys ^ 0x80000000; down: keys ^ 0x7fffffff; both ranked
 * ascending, stably.  lens == NULL (or m <= 1) is the unsegmented rank.
 * work needs 3*n int32. */
void orc_rank_z(int32_t *rank, const int32_t *s, int64_t n,
                const int32_t *lens, int64_t m, int up, int32_t *work)
{
    uint32_t mask = up ? 0x80000000u : 0x7fffffffu;
    int32_t *order = work, *tmp = work + n;
    uint32_t *key = (uint32_t *)(work + 2 * n);
    for (int64_t k = 0; k < n; k++) { order[k] = (int32_t)k; key[k] = (uint32_t)s[k] ^ mask; }
    orc_field_rank(order, key, tmp, n);
    orc_rank_finish(rank, order, tmp, key, lens, m, n);
}

/* Double rank, ref rank.c:234-315 (double_rank_sort) with the
 * little-endian word order the reference selects for x86 (rank.c:225-231):
 * low 32-bit word first, then the high word; with the sign bit s of the
 * double, a field is ~t when s is set, else t (low word) or t ^ 0x80000000
 * (high word); down ranks use the complement of each field.
 * work needs 3*n int32. */
void orc_rank_d(int32_t *rank, const double *s, int64_t n,
                const int32_t *lens, int64_t m, int up, int32_t *work)
{
    int32_t *order = work, *tmp = work + n;
    uint32_t *fld = (uint32_t *)(work + 2 * n);
    for (int64_t k = 0; k < n; k++) order[k] = (int32_t)k;
    for (int word = 0; word < 2; word++) {
        for (int64_t k = 0; k < n; k++) {
            uint64_t bits;
            memcpy(&bits, &s[k], 8);
            uint32_t t = word == 0 ? (uint32_t)bits : (uint32_t)(bits >> 32);
            int neg = (bits >> 63) != 0;
            uint32_t f = neg ? ~t : (word == 0 ? t : t ^ 0x80000000u);
            fld[k] = up ? f : ~f;
        }
        orc_field_rank(order, fld, tmp, n);
    }
    orc_rank_finish(rank, order, tmp, fld, lens, m, n);
}

/* ------------------------------------------------------------------ */
/* Elementwise operations                                              */
/* ------------------------------------------------------------------ */

/* Binary elementwise codes.  Semantics: ref elwise.c:179-261 with the
 * operator macros of defins.h:61-113. */
enum {
    EW_MAX = 0, EW_MIN, EW_ADD, EW_SUB, EW_MUL, EW_DIV,
    EW_GRT, EW_LES, EW_GEQ, EW_LEQ, EW_EQL, EW_NEQ,
    EW_LSH, EW_RSH, EW_MOD, EW_AND, EW_IOR, EW_XOR
};

/* z inputs: int32 in, result int32 (arith/bitwise) or cvl_bool (compare).
 * Division and modulo follow C (truncation); shifts are C shifts. */
int orc_ew2_z(int op, int32_t *d, const int32_t *a, const int32_t *b, int64_t n)
{
    for (int64_t k = 0; k < n; k++) {
        int32_t x = a[k], y = b[k], r;
        switch (op) {
        case EW_MAX: r = x > y ? x : y; break;
        case EW_MIN: r = x < y ? x : y; break;
        case EW_ADD: r = (int32_t)((uint32_t)x + (uint32_t)y); break;
        case EW_SUB: r = (int32_t)((uint32_t)x - (uint32_t)y); break;
        case EW_MUL: r = (int32_t)((uint32_t)x * (uint32_t)y); break;
        case EW_DIV: r = x / y; break;
        case EW_GRT: r = x > y; break;
        case EW_LES: r = x < y; break;
        case EW_GEQ: r = x >= y; break;
        case EW_LEQ: r = x <= y; break;
        case EW_EQL: r = x == y; break;
        case EW_NEQ: r = x != y; break;
        case EW_LSH: r = (int32_t)((uint32_t)x << y); break;
        case EW_RSH: r = x >> y; break;
        case EW_MOD: r = x % y; break;
        case EW_AND: r = x & y; break;
        case EW_IOR: r = x | y; break;
        case EW_XOR: r = x ^ y; break;
        default: return -1;
        }
        d[k] = r;
    }
    return 0;
}

/* d inputs: doubles; arithmetic gives doubles (written to d64), compares
 * give cvl_bool (written to d32). */
int orc_ew2_d(int op, double *d64, int32_t *d32, const double *a,
              const double *b, int64_t n)
{
    for (int64_t k = 0; k < n; k++) {
        double x = a[k], y = b[k];
        switch (op) {
        case EW_MAX: d64[k] = x > y ? x : y; break;
        case EW_MIN: d64[k] = x < y ? x : y; break;
        case EW_ADD: d64[k] = x + y; break;
        case EW_SUB: d64[k] = x - y; break;
        case EW_MUL: d64[k] = x * y; break;
        case EW_DIV: d64[k] = x / y; break;
        case EW_GRT: d32[k] = x > y; break;
        case EW_LES: d32[k] = x < y; break;
        case EW_GEQ: d32[k] = x >= y; break;
        case EW_LEQ: d32[k] = x <= y; break;
        case EW_EQL: d32[k] = x == y; break;
        case EW_NEQ: d32[k] = x != y; break;
        default: return -1;
        }
    }
    return 0;
}

/* b inputs: logical ops and compares on cvl_bool (elwise.c:242-255). */
int orc_ew2_b(int op, int32_t *d, const int32_t *a, const int32_t *b, int64_t n)
{
    for (int64_t k = 0; k < n; k++) {
        int32_t x = a[k], y = b[k], r;
        switch (op) {
        case EW_EQL: r = x == y; break;
        case EW_NEQ: r = x != y; break;
        case EW_AND: r = x && y; break;
        case EW_IOR: r = x || y; break;
        case EW_XOR: r = x ^ y; break;   /* elwise.c:252: bitwise ^ */
        default: return -1;
        }
        d[k] = r;
    }
    return 0;
}

/* Unary codes (elwise.c:251-289). */
enum {
    EU_NOTB = 0, EU_NOTZ, EU_INT_D, EU_INT_B, EU_DBL_Z, EU_BOO_Z,
    EU_FLR, EU_CEI, EU_TRN, EU_ROU, EU_EXP, EU_LOG, EU_SQT, EU_SIN, EU_COS,
    EU_TAN, EU_ASN, EU_ACS, EU_ATN, EU_SNH, EU_CSH, EU_TNH
};

/* Unary ops: s and d are int32 or double as the op implies. */
int orc_ew1(int op, void *d, const void *s, int64_t n)
{
    const int32_t *zi = (const int32_t *)s; const double *di = (const double *)s;
    int32_t *zo = (int32_t *)d; double *dd = (double *)d;
    for (int64_t k = 0; k < n; k++) {
        switch (op) {
        case EU_NOTB: zo[k] = !zi[k]; break;
        case EU_NOTZ: zo[k] = ~zi[k]; break;
        case EU_INT_D: zo[k] = (int32_t)di[k]; break;
        case EU_INT_B: zo[k] = zi[k]; break;
        case EU_DBL_Z: dd[k] = (double)zi[k]; break;
        case EU_BOO_Z: zo[k] = zi[k] != 0; break;
        case EU_FLR: zo[k] = (int32_t)floor(di[k]); break;
        case EU_CEI: zo[k] = (int32_t)ceil(di[k]); break;
        case EU_TRN: zo[k] = (int32_t)di[k]; break;
        case EU_ROU: zo[k] = (int32_t)(di[k] + 0.5); break;   /* defins.h:121 */
        case EU_EXP: dd[k] = exp(di[k]); break;
        case EU_LOG: dd[k] = log(di[k]); break;
        case EU_SQT: dd[k] = sqrt(di[k]); break;
        case EU_SIN: dd[k] = sin(di[k]); break;
        case EU_COS: dd[k] = cos(di[k]); break;
        case EU_TAN: dd[k] = tan(di[k]); break;
        case EU_ASN: dd[k] = asin(di[k]); break;
        case EU_ACS: dd[k] = acos(di[k]); break;
        case EU_ATN: dd[k] = atan(di[k]); break;
        case EU_SNH: dd[k] = sinh(di[k]); break;
        case EU_CSH: dd[k] = cosh(di[k]); break;
        case EU_TNH: dd[k] = tanh(di[k]); break;
        default: return -1;
        }
    }
    return 0;
}

/* select: d[k] = f[k] ? a[k] : b[k] (elwise.c:165-177, 246-248). */
void orc_sel(int ty, void *d, const int32_t *f, const void *a, const void *b,
             int64_t n)
{
    size_t w = W_OF(ty);
    for (int64_t k = 0; k < n; k++)
        memcpy((char *)d + k * w, f[k] ? (const char *)a + k * w
                                       : (const char *)b + k * w, w);
}

/* ------------------------------------------------------------------ */
/* RAND: lrand48() % s[k]                                              */
/* ------------------------------------------------------------------ */

/* The 48-bit linear congruential generator behind lrand48 (POSIX:
 * X' = (0x5DEECE66D * X + 0xB) mod 2^48, result = X' >> 17).  RAND is
 * lrand48() % s[k] for each element in order (ref elwise.c:211-236,
 * defins.h:86-97).  *state is the 48-bit X, updated in place. */
int32_t orc_lrand48_next(uint64_t *state)
{
    uint64_t x = (*state * 0x5DEECE66DULL + 0xBULL) & ((1ULL << 48) - 1);
    *state = x;
    return (int32_t)(x >> 17);
}

void orc_rnd_wuz(int32_t *d, const int32_t *s, int64_t n, uint64_t *state)
{
    for (int64_t k = 0; k < n; k++) d[k] = orc_lrand48_next(state) % s[k];
}

/* ------------------------------------------------------------------ */
/* Synthetic inputs: splitmix64                                        */
/* ------------------------------------------------------------------ */

/* splitmix64 (Steele, Lea, Flood 2014): element k is the mix of
 * seed + (k+1) * 0x9E3779B97F4A7C15.  The HIP backend computes the same
 * counter-based stream, so host and device inputs are identical. */
static inline uint64_t orc_sm64(uint64_t seed, uint64_t k)
{
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void orc_splitmix_u32(uint32_t *d, int64_t n, uint64_t seed)
{
    for (int64_t k = 0; k < n; k++) d[k] = (uint32_t)(orc_sm64(seed, (uint64_t)k) >> 32);
}

/* Uniform doubles in [lo, hi): 53 random bits. */
void orc_splitmix_f64(double *d, int64_t n, uint64_t seed, double lo, double hi)
{
    for (int64_t k = 0; k < n; k++) {
        double u = (double)(orc_sm64(seed, (uint64_t)k) >> 11) * 0x1.0p-53;
        d[k] = lo + (hi - lo) * u;
    }
}
