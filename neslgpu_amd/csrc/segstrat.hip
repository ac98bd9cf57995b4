// segstrat.hip -- schedules for segmented scans and reductions, and the
// descriptor statistics the dispatcher reads.
//
// A segmented scan or reduction (ref nesl/cvl/serial/vprims.c:62-201) can
// be scheduled several ways; all of them compute the serial semantics:
//   arm 0  flat, 16-row tiles   decoupled look-back over the flat vector
//   arm 1  flat, 8-row tiles
//   arm 2  thread per segment   one lane walks a segment sequentially
//   arm 3  wave per segment     a wave scans 64x4 (int) / 64x2 (double)
//                               elements per step with a DPP scan, carry
//                               across steps
//   arm 4  block per segment    256 threads, 4 waves per step, carry
//   arm 5  block per segment    1024 threads
// For exact operators (integers, min, max) every arm gives bit-identical
// results.  For double + the association differs between arms (each arm is
// deterministic); all stay within the compensated-sum tolerance.
//
// Statistics (cvlhip_segd_features): element and segment fractions per
// log2-length class, log2 n, log2 m, log2 mean length -- computed on the
// device from the lengths.
#include "cvl_hip.hpp"
#include "segd.hpp"
#include "scan.hpp"
#include "../../include/cvl.h"
#include "../../include/cvl_hip_ext.h"

namespace cvlhip {

constexpr int kArms = 6;
constexpr int kLenBins = 21;   // length classes [2^i, 2^(i+1)), i = 0..20 (the last is open)
constexpr int kFeatures = 3 + 2 * kLenBins + 2;

// ------------------------------------------------------------------ thread per segment
template <class Op, bool REDUCE>
__global__ void k_seg_thread(const typename Op::T *__restrict__ in, typename Op::T *__restrict__ out, SegdView sv) {
    using T = typename Op::T;
    const Op op;
    const T cid = Op::cvl_id();
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < sv.m; g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t lo = sv.offs[g], hi = lo + sv.lens[g];
        T acc = Op::neutral();
        for (int64_t k = lo; k < hi; k++) {
            const T x = load_val<Op>(in[k]);
            if constexpr (!REDUCE) out[k] = op(cid, acc);
            acc = op(acc, x);
        }
        if constexpr (REDUCE) out[g] = op(cid, acc);
    }
}

// ------------------------------------------------------------------ wave per segment
template <class Op, bool REDUCE>
__global__ __launch_bounds__(256) void k_seg_wave(const typename Op::T *__restrict__ in,
                                                  typename Op::T *__restrict__ out, SegdView sv) {
    using T = typename Op::T;
    constexpr int V = 16 / (int)sizeof(T);
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
#pragma unroll
            for (int k = 0; k < V; k++) x[k] = b + k < hi ? load_val<Op>(in[b + k]) : Op::neutral();
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

// ------------------------------------------------------------------ block per segment
template <class Op, bool REDUCE, int THREADS>
__global__ __launch_bounds__(THREADS) void k_seg_block(const typename Op::T *__restrict__ in,
                                                       typename Op::T *__restrict__ out, SegdView sv) {
    using T = typename Op::T;
    constexpr int V = 16 / (int)sizeof(T);
    constexpr int W = THREADS / kWave;
    constexpr int STEP = THREADS * V;
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
#pragma unroll
            for (int k = 0; k < V; k++) x[k] = b + k < hi ? load_val<Op>(in[b + k]) : Op::neutral();
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

// ------------------------------------------------------------------ statistics
__global__ void k_seg_features(const int *__restrict__ lens, int64_t m, unsigned long long *__restrict__ acc) {
    // acc[0..20]: elements per length class, acc[21..41]: segments per class
    __shared__ unsigned long long s[2 * kLenBins];
    for (int i = threadIdx.x; i < 2 * kLenBins; i += blockDim.x) s[i] = 0;
    __syncthreads();
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < m; g += (int64_t)gridDim.x * blockDim.x) {
        const unsigned L = (unsigned)lens[g];
        if (!L) continue;
        int c = 31 - __clz(L);
        if (c >= kLenBins) c = kLenBins - 1;
        atomicAdd(&s[c], (unsigned long long)L);
        atomicAdd(&s[kLenBins + c], 1ull);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * kLenBins; i += blockDim.x)
        if (s[i]) atomicAdd(acc + i, s[i]);
}

template <class Op>
static void run_arm(bool reduce, int arm, void *d, const void *s, const void *sd, int64_t n, int64_t m,
                    void *scratch) {
    using T = typename Op::T;
    if (m <= 0) return;
    SegdView sv(sd, n, m);
    int g;
    switch (arm) {
    case 0:
        if (reduce) reduce_seg_r16<Op>(d, s, sd, n, m, scratch);
        else scan_seg_r16<Op>(d, s, sd, n, m, scratch);
        return;
    case 1:
        if (reduce) reduce_seg_r8<Op>(d, s, sd, n, m, scratch);
        else scan_seg_r8<Op>(d, s, sd, n, m, scratch);
        return;
    case 2:
        g = grid_for(m, 256);
        if (g > 65535) g = 65535;
        if (reduce) k_seg_thread<Op, true><<<g, 256, 0, g_stream>>>((const T *)s, (T *)d, sv);
        else k_seg_thread<Op, false><<<g, 256, 0, g_stream>>>((const T *)s, (T *)d, sv);
        break;
    case 3:
        g = grid_for(m, 256 / kWave);
        if (g > 65535) g = 65535;
        if (reduce) k_seg_wave<Op, true><<<g, 256, 0, g_stream>>>((const T *)s, (T *)d, sv);
        else k_seg_wave<Op, false><<<g, 256, 0, g_stream>>>((const T *)s, (T *)d, sv);
        break;
    case 4:
        g = (int)(m < 65535 ? m : 65535);
        if (reduce) k_seg_block<Op, true, 256><<<g, 256, 0, g_stream>>>((const T *)s, (T *)d, sv);
        else k_seg_block<Op, false, 256><<<g, 256, 0, g_stream>>>((const T *)s, (T *)d, sv);
        break;
    default:
        g = (int)(m < 65535 ? m : 65535);
        if (reduce) k_seg_block<Op, true, 1024><<<g, 1024, 0, g_stream>>>((const T *)s, (T *)d, sv);
        else k_seg_block<Op, false, 1024><<<g, 1024, 0, g_stream>>>((const T *)s, (T *)d, sv);
        break;
    }
    LAUNCHCHECK();
}

template <class Op>
static void run_arm_checked(bool reduce, int arm, void *d, const void *s, const void *sd, int64_t n, int64_t m,
                            void *scratch) {
    if (arm < 0 || arm >= kArms) { usage_error("segmented schedule %d does not exist", arm); return; }
    run_arm<Op>(reduce, arm, d, s, sd, n, m, scratch);
}

static void dispatch(int op, int ty, bool reduce, int arm, void *d, const void *s, const void *sd, int64_t n,
                     int64_t m, void *scratch) {
    // op: 0 add, 1 mul, 2 min, 3 max; ty: 0 int, 1 double
    if (ty == 0) {
        switch (op) {
        case 0: run_arm_checked<AddZ>(reduce, arm, d, s, sd, n, m, scratch); return;
        case 1: run_arm_checked<MulZ>(reduce, arm, d, s, sd, n, m, scratch); return;
        case 2: run_arm_checked<MinZ>(reduce, arm, d, s, sd, n, m, scratch); return;
        case 3: run_arm_checked<MaxZ>(reduce, arm, d, s, sd, n, m, scratch); return;
        }
    } else if (ty == 1) {
        switch (op) {
        case 0: run_arm_checked<AddD>(reduce, arm, d, s, sd, n, m, scratch); return;
        case 1: run_arm_checked<MulD>(reduce, arm, d, s, sd, n, m, scratch); return;
        case 2: run_arm_checked<MinD>(reduce, arm, d, s, sd, n, m, scratch); return;
        case 3: run_arm_checked<MaxD>(reduce, arm, d, s, sd, n, m, scratch); return;
        }
    }
    usage_error("segmented op %d / type %d is not scheduled here", op, ty);
}

static unsigned long long *g_feat_dev = nullptr;
static unsigned long long *g_feat_host = nullptr;

}  // namespace cvlhip

using namespace cvlhip;

extern "C" {

int cvlhip_seg_arms(void) { return kArms; }
int cvlhip_seg_features_count(void) { return kFeatures; }

void cvlhip_seg_scan(int op, int ty, int arm, vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch) {
    dispatch(op, ty, false, arm, d, s, sd, n, m, scratch);
}

void cvlhip_seg_reduce(int op, int ty, int arm, vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch) {
    dispatch(op, ty, true, arm, d, s, sd, n, m, scratch);
}

int cvlhip_seg_scratch(int n, int m) {
    (void)m;
    int64_t a = scan_scratch_bytes<AddD>(n) * 2, b = kReduceScratchBytes;
    const int64_t c = seg_reduce_scratch_bytes(n, 8, 8);   // the 8-row arm has the most tiles
    if (c > a) a = c;
    return units_for_bytes(a > b ? a : b);
}

// Features of a descriptor, written to host memory `out` (kFeatures floats):
// [0] log2(n)/32  [1] log2(m)/32  [2] log2(n/m)/20
// [3..23] fraction of elements in length class i
// [24..44] fraction of segments in length class i
// [45] is_double  [46] is_reduce  (set by the caller; written as 0 here)
void cvlhip_segd_features(vec_p sd, int n, int m, float *out) {
    if (!g_feat_dev) {
        HIPCHECK(hipMalloc(&g_feat_dev, 8 * 2 * kLenBins));
        HIPCHECK(hipHostMalloc(&g_feat_host, 8 * 2 * kLenBins, hipHostMallocDefault));
    }
    fill_u32(g_feat_dev, 0, 2 * 2 * kLenBins);
    if (m > 0) {
        SegdView sv(sd, n, m);
        int g = grid_for(m, 256);
        if (g > 1024) g = 1024;
        k_seg_features<<<g, 256, 0, g_stream>>>(sv.lens, m, g_feat_dev);
        LAUNCHCHECK();
    }
    HIPCHECK(hipMemcpyAsync(g_feat_host, g_feat_dev, 8 * 2 * kLenBins, hipMemcpyDeviceToHost, g_stream));
    HIPCHECK(hipStreamSynchronize(g_stream));
    const double nn = n > 0 ? n : 1, mm = m > 0 ? m : 1;
    out[0] = (float)(log2(nn) / 32.0);
    out[1] = (float)(log2(mm) / 32.0);
    out[2] = (float)(log2(nn / mm > 1 ? nn / mm : 1) / 20.0);
    for (int i = 0; i < kLenBins; i++) {
        out[3 + i] = (float)(g_feat_host[i] / nn);
        out[3 + kLenBins + i] = (float)(g_feat_host[kLenBins + i] / mm);
    }
    out[3 + 2 * kLenBins] = 0.f;
    out[4 + 2 * kLenBins] = 0.f;
}

}  // extern "C"
