// runtime.hip -- library state, error handling, and the CVL facility
// functions: sizes, memory, moves, host<->device copies, timers, RNG seed.
//
// Facility semantics follow the serial CVL (ref nesl/cvl/serial/facilt.c);
// memory is device memory: vec_p is a device pointer, one CVL unit is
// 8 bytes (the int/float/double union of facilt.c:92-93), and alo_foz is a
// hipMalloc arena the caller sub-allocates (ref nesl/vcode/vstack.c:156-160).
#include "cvl_hip.hpp"
#include "../../include/cvl.h"
#include "../../include/cvl_hip_ext.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

namespace cvlhip {

hipStream_t g_stream = nullptr;
static int g_fatal = 1;
static char g_errbuf[512];
static int g_err = 0;
static unsigned *g_dev_err = nullptr;

void fatal(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "cvl-hip: ");
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    va_end(ap);
    exit(1);
}

void usage_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_errbuf, sizeof g_errbuf, fmt, ap);
    va_end(ap);
    g_err = 1;
    if (g_fatal) fatal("%s", g_errbuf);
}

void hip_fail(hipError_t e, const char *expr, const char *file, int line) {
    snprintf(g_errbuf, sizeof g_errbuf, "%s failed: %s (%s:%d)", expr, hipGetErrorString(e), file, line);
    g_err = 2;
    if (g_fatal) fatal("%s", g_errbuf);
}

// Allocated and cleared by the first call, which the library makes in
// cvlhip_set_stream, CVL_init, alo_foz and cvlhip_capture_begin: hosts
// start with one of those, so the clear is never recorded into a captured
// graph.  The clear is synchronous (hipMemset), before any kernel can read
// the word.
unsigned *device_error_word() {
    if (!g_dev_err) {
        HIPCHECK(hipMalloc(&g_dev_err, 256));
        HIPCHECK(hipMemset(g_dev_err, 0, 256));
    }
    return g_dev_err;
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t v, int64_t count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = v;
}

static Mailbox *g_mb_host = nullptr;
static Mailbox *g_mb_dev = nullptr;
static unsigned long long g_mb_seq = 0;

Mailbox *mailbox_device() {
    if (!g_mb_host) {
        HIPCHECK(hipHostMalloc((void **)&g_mb_host, 256,
                               hipHostMallocMapped | hipHostMallocCoherent | hipHostMallocPortable));
        g_mb_host->seq = 0;
        g_mb_host->value = 0;
        HIPCHECK(hipHostGetDevicePointer((void **)&g_mb_dev, g_mb_host, 0));
    }
    return g_mb_dev;
}

unsigned long long mailbox_next() { return ++g_mb_seq; }

// Spins on the sequence number for up to 2 ms, then waits for the stream
// (which reports a failed kernel) and looks once more.
unsigned long long mailbox_wait(unsigned long long seq) {
    const struct timespec t0 = [] { struct timespec t; clock_gettime(CLOCK_MONOTONIC, &t); return t; }();
    for (unsigned it = 1;; it++) {
        if (__atomic_load_n(&g_mb_host->seq, __ATOMIC_ACQUIRE) == seq)
            return __atomic_load_n(&g_mb_host->value, __ATOMIC_RELAXED);
        if ((it & 1023) == 0) {
            struct timespec t;
            clock_gettime(CLOCK_MONOTONIC, &t);
            if ((t.tv_sec - t0.tv_sec) * 1000000000LL + (t.tv_nsec - t0.tv_nsec) > 2000000LL) {
                HIPCHECK(hipStreamSynchronize(g_stream));
                if (__atomic_load_n(&g_mb_host->seq, __ATOMIC_ACQUIRE) == seq)
                    return __atomic_load_n(&g_mb_host->value, __ATOMIC_RELAXED);
                fatal("a kernel finished without posting its scalar result (sequence %llu)", seq);
            }
        }
        __builtin_ia32_pause();
    }
}

void fill_u32(void *p, uint32_t v, int64_t count) {
    if (count <= 0) return;
    int g = grid_for(count, kBlock * 4);
    if (g > 8192) g = 8192;
    k_fill_u32<<<g, kBlock, 0, g_stream>>>((uint32_t *)p, v, count);
    LAUNCHCHECK();
}

// Up to three regions cleared in one launch (element i < n1 of the first,
// else of the second, else of the third): a launch saved where they are
// small.
__global__ void k_zero3_u32(uint32_t *__restrict__ p1, int64_t n1, uint32_t *__restrict__ p2, int64_t n2,
                            uint32_t *__restrict__ p3, int64_t n3) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n1 + n2 + n3; i += stride) {
        if (i < n1) p1[i] = 0u;
        else if (i < n1 + n2) p2[i - n1] = 0u;
        else p3[i - n1 - n2] = 0u;
    }
}

void zero3_u32(void *p1, int64_t n1, void *p2, int64_t n2, void *p3, int64_t n3) {
    if (n1 < 0 || !p1) n1 = 0;
    if (n2 < 0 || !p2) n2 = 0;
    if (n3 < 0 || !p3) n3 = 0;
    if (n1 + n2 + n3 <= 0) return;
    int g = grid_for(n1 + n2 + n3, kBlock * 4);
    if (g > 8192) g = 8192;
    k_zero3_u32<<<g, kBlock, 0, g_stream>>>((uint32_t *)p1, n1, (uint32_t *)p2, n2, (uint32_t *)p3, n3);
    LAUNCHCHECK();
}

void zero2_u32(void *p1, int64_t n1, void *p2, int64_t n2) { zero3_u32(p1, n1, p2, n2, nullptr, 0); }

// Overlap-safe move of `bytes` bytes (ref facilt.c:161-183 walks forward
// when dest < src and backward otherwise).  Each block copies a chunk that
// does not overlap the chunks still to be read by later blocks in the same
// launch, so large overlapping moves are split into launches of
// non-overlapping pieces.
__global__ void k_copy_bytes16(const uint4 *__restrict__ s, uint4 *__restrict__ d, int64_t n16) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) d[i] = s[i];
}

static void copy_nonoverlap(void *d, const void *s, int64_t bytes) {
    if (bytes <= 0) return;
    HIPCHECK(hipMemcpyAsync(d, s, (size_t)bytes, hipMemcpyDeviceToDevice, g_stream));
}

void move_units(void *d, const void *s, int64_t units) {
    if (d == s || units <= 0) return;
    int64_t bytes = units * kUnit;
    char *dc = (char *)d;
    const char *sc = (const char *)s;
    int64_t dist = dc > sc ? dc - sc : sc - dc;
    if (dist >= bytes) { copy_nonoverlap(d, s, bytes); return; }
    // Overlapping: copy in pieces of `dist` bytes, in the direction that
    // never overwrites bytes not yet read.  Pieces are stream-ordered.
    if (dc < sc) {
        for (int64_t off = 0; off < bytes; off += dist) {
            int64_t len = bytes - off < dist ? bytes - off : dist;
            copy_nonoverlap(dc + off, sc + off, len);
        }
    } else {
        for (int64_t end = bytes; end > 0; end -= dist) {
            int64_t len = end < dist ? end : dist;
            copy_nonoverlap(dc + end - len, sc + end - len, len);
        }
    }
}

}  // namespace cvlhip

using namespace cvlhip;

// ------------------------------------------------------------------ extension API
extern "C" {

void cvlhip_set_stream(void *s) {
    device_error_word();
    g_stream = (hipStream_t)s;
}
void *cvlhip_get_stream(void) { return (void *)g_stream; }
void cvlhip_set_error_mode(int fatal_on_error) { g_fatal = fatal_on_error; }

int cvlhip_last_error(char *buf, int len) {
    int e = g_err;
    if (buf && len > 0) {
        strncpy(buf, g_err ? g_errbuf : "", (size_t)len - 1);
        buf[len - 1] = 0;
    }
    g_err = 0;
    return e;
}

unsigned cvlhip_check_device_errors(void) {
    unsigned h[2] = {0, 0};
    unsigned *w = device_error_word();
    unsigned *w2 = noalloc_error_word();        // the no-allocation entry's, once it ran here
    HIPCHECK(hipMemcpyAsync(h, w, 4, hipMemcpyDeviceToHost, g_stream));
    if (w2) HIPCHECK(hipMemcpyAsync(h + 1, w2, 4, hipMemcpyDeviceToHost, g_stream));
    HIPCHECK(hipStreamSynchronize(g_stream));
    if (h[0]) HIPCHECK(hipMemsetAsync(w, 0, 4, g_stream));
    if (h[1]) HIPCHECK(hipMemsetAsync(w2, 0, 4, g_stream));
    return h[0] | h[1];
}

void cvlhip_synchronize(void) { HIPCHECK(hipStreamSynchronize(g_stream)); }

void *cvlhip_device_error_word(void) { return (void *)device_error_word(); }

// ---- HIP graphs of CVL call sequences
void *cvlhip_stream_create(void) {
    hipStream_t s = nullptr;
    HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return (void *)s;
}

int cvlhip_capture_begin(void) {
    if (g_stream == nullptr) return -1;                // the null stream cannot be captured
    device_error_word();                               // allocated outside the capture
    return hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal) == hipSuccess ? 0 : -1;
}

int cvlhip_capture_end(void **exec) {
    hipGraph_t g = nullptr;
    *exec = nullptr;
    if (hipStreamEndCapture(g_stream, &g) != hipSuccess || g == nullptr) {
        (void)hipGetLastError();
        return -1;
    }
    hipGraphExec_t e = nullptr;
    hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (r != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    *exec = (void *)e;
    return 0;
}

int cvlhip_graph_launch(void *exec) {
    return hipGraphLaunch((hipGraphExec_t)exec, g_stream) == hipSuccess ? 0 : -1;
}

void cvlhip_graph_destroy(void *exec) {
    if (exec) (void)hipGraphExecDestroy((hipGraphExec_t)exec);
}


// ------------------------------------------------------------------ CVL facilities
void CVL_init(char *fused_library) {
    (void)fused_library;   // fused libraries are loaded by the interpreter
    device_error_word();
}

// Sizes in CVL units (ref facilt.c:95-116).  The segment descriptor of this
// backend is larger than the serial one: see segd.hip.
int siz_fob(int length) { return units_for_bytes((int64_t)length * 4); }
int siz_foz(int length) { return units_for_bytes((int64_t)length * 4); }
int siz_fod(int length) { return units_for_bytes((int64_t)length * 8); }

vec_p alo_foz(int size) {
    void *p = nullptr;
    if (size < 0) return nullptr;
    device_error_word();
    if (hipMalloc(&p, (size_t)(size > 0 ? size : 1) * kUnit) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;   // the caller reports failure (facilt.c:149-153)
    }
    return p;
}

void fre_fov(vec_p pointer) {
    if (pointer) HIPCHECK(hipFree(pointer));
}

void mov_fov(vec_p d, vec_p s, int size, vec_p scratch) {
    (void)scratch;
    move_units(d, s, size);
}
int mov_fov_scratch(int size) { (void)size; return 0; }

vec_p add_fov(vec_p v, int a) { return (vec_p)((char *)v + (int64_t)a * kUnit); }

int sub_fov(vec_p v1, vec_p v2) {
    int64_t d = ((char *)v1 - (char *)v2) / kUnit;
    return (int)(d < 0 ? -d : d);
}

cvl_bool eql_fov(vec_p v1, vec_p v2) { return v1 == v2; }

int cmp_fov(vec_p v1, vec_p v2) { return (int)(((char *)v1 - (char *)v2) / kUnit); }

// Host <-> device copies (ref facilt.c:241-270).  The CVL call returns with
// the host buffer filled (v2c) or already consumed (c2v).
#define CVLHIP_V2C(name, ctype)                                                   \
    void name(ctype *d, vec_p s, int len, vec_p scratch) {                        \
        (void)scratch;                                                            \
        if (len <= 0) return;                                                     \
        HIPCHECK(hipMemcpyAsync(d, s, sizeof(ctype) * (size_t)len, hipMemcpyDeviceToHost, g_stream)); \
        HIPCHECK(hipStreamSynchronize(g_stream));                                 \
    }                                                                             \
    int name##_scratch(int len) { (void)len; return 0; }                          \
    unsigned int name##_inplace(void) { return INPLACE_NONE; }
#define CVLHIP_C2V(name, ctype)                                                   \
    void name(vec_p d, ctype *s, int len, vec_p scratch) {                        \
        (void)scratch;                                                            \
        if (len <= 0) return;                                                     \
        HIPCHECK(hipMemcpyAsync(d, s, sizeof(ctype) * (size_t)len, hipMemcpyHostToDevice, g_stream)); \
        HIPCHECK(hipStreamSynchronize(g_stream));                                 \
    }                                                                             \
    int name##_scratch(int len) { (void)len; return 0; }                          \
    unsigned int name##_inplace(void) { return INPLACE_NONE; }

CVLHIP_V2C(v2c_fuz, int)
CVLHIP_V2C(v2c_fub, cvl_bool)
CVLHIP_V2C(v2c_fud, double)
CVLHIP_C2V(c2v_fuz, int)
CVLHIP_C2V(c2v_fub, cvl_bool)
CVLHIP_C2V(c2v_fud, double)

// Timers (ref facilt.c:35-85): wall clock, as the serial CVL on Linux.
// Device work issued before the call is finished first, so that
// START_TIMER/STOP_TIMER bracket completed GPU work.
void tgt_fos(cvl_timer_t *t) {
    HIPCHECK(hipStreamSynchronize(g_stream));
    clock_gettime(CLOCK_REALTIME, t);
}

double tdf_fos(cvl_timer_t *t1, cvl_timer_t *t2) {
    return (double)(t1->tv_sec - t2->tv_sec) + 1.0e-9 * (double)(t1->tv_nsec - t2->tv_nsec);
}

}  // extern "C"
