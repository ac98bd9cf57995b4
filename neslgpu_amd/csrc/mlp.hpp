// mlp.hpp -- shape of the learned dispatcher's network and its one-wave
// forward pass used when a segment descriptor is built.  No counterpart in
// the reference.
//
// Network: x (kF features) -> ReLU(W1 x + b1) (kH) -> ReLU(W2 h1 + b2) (kH)
//          -> W3 h2 + b3 (kA outputs, one predicted log-time per schedule).
// Parameters live in one flat array in the order W1 b1 W2 b2 W3 b3
// (row-major W[out][in]).
#pragma once

#include "cvl_hip.hpp"
#include "segd.hpp"

namespace cvlhip {

constexpr int kF = 3 + 2 * kLenClasses + 2, kH = 32, kA = 6;
constexpr int kW1 = 0, kB1 = kW1 + kH * kF, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH,
              kB3 = kW3 + kA * kH, kParams = kB3 + kA;
static_assert(kH <= kWave && kA <= kWave && kF <= kWave, "one wave holds a layer");

// Feature i of a descriptor of n elements in m segments whose length-class
// histogram is elems[]/segs[]:
//   [0] log2(n)/32   [1] log2(m)/32   [2] log2(max(n/m, 1))/20
//   [3 + c]      fraction of the elements in segments of class c
//   [3 + 21 + c] fraction of the segments in class c
//   [45] 1 for 8-byte values   [46] 1 for a reduction
__device__ __forceinline__ float sched_feature(int i, int64_t n, int64_t m, const uint32_t *elems,
                                               const uint32_t *segs, bool dbl, bool red) {
    const float nn = n > 0 ? (float)n : 1.f, mm = m > 0 ? (float)m : 1.f;
    if (i == 0) return log2f(nn) / 32.f;
    if (i == 1) return log2f(mm) / 32.f;
    if (i == 2) return log2f(nn / mm > 1.f ? nn / mm : 1.f) / 20.f;
    if (i < 3 + kLenClasses) return (float)elems[i - 3] / nn;
    if (i < 3 + 2 * kLenClasses) return (float)segs[i - 3 - kLenClasses] / mm;
    if (i == 3 + 2 * kLenClasses) return dbl ? 1.f : 0.f;
    if (i == 4 + 2 * kLenClasses) return red ? 1.f : 0.f;
    return 0.f;
}

// Schedules a descriptor admits whatever the network predicts.  The flat
// schedules (0-2) take any descriptor.  The per-segment ones (3 thread, 4
// wave, 5 workgroup) give a segment one lane, wave or workgroup: one long
// segment serialises them and a few segments leave the chip idle, so each
// is admitted only when the longest segment is below 2^(kArmMaxClass + 1)
// and there are at least 2^kArmMinLogM segments.  cmax: the highest
// length class holding a segment (-1: none).  Bit a set: schedule a admitted.
constexpr int kArmMaxClass[3] = {5, 12, 15};
constexpr int kArmMinLogM[3] = {14, 11, 9};
__host__ __device__ inline unsigned sched_admitted(int cmax, int64_t m) {
    unsigned a = 7u;
    for (int k = 0; k < 3; k++)
        if (cmax <= kArmMaxClass[k] && m >= ((int64_t)1 << kArmMinLogM[k])) a |= 1u << (3 + k);
    return a;
}
__host__ __device__ inline int max_len_class(const uint32_t *segs) {
    int c = kLenClasses - 1;
    while (c >= 0 && segs[c] == 0) c--;
    return c;
}

// Forward pass of one feature vector by one whole wave (lane i holds
// feature i); every lane returns the index of the smallest prediction among
// the admitted schedules (the first one on ties).  Values move between lanes with shuffles only, and
// sums run in index order, so the same parameters and features always give
// the same pick.
__device__ __forceinline__ int mlp_pick_wave(const float *__restrict__ P, float feature_of_lane, unsigned admitted) {
    const int lane = lane_id();
    const int j1 = lane < kH ? lane : kH - 1;       // lanes past a layer's width compute a copy
    float a = P[kB1 + j1];
    for (int i = 0; i < kF; i++) a += P[kW1 + j1 * kF + i] * shfl(feature_of_lane, i);
    const float h1 = a > 0.f ? a : 0.f;
    a = P[kB2 + j1];
    for (int i = 0; i < kH; i++) a += P[kW2 + j1 * kH + i] * shfl(h1, i);
    const float h2 = a > 0.f ? a : 0.f;
    const int j3 = lane < kA ? lane : kA - 1;
    a = P[kB3 + j3];
    for (int i = 0; i < kH; i++) a += P[kW3 + j3 * kH + i] * shfl(h2, i);
    int best = 0;                                   // flat8 is always admitted
    float by = shfl(a, 0);
    for (int k = 1; k < kA; k++) {
        const float v = shfl(a, k);
        if (((admitted >> k) & 1u) && v < by) { by = v; best = k; }
    }
    return best;
}

}  // namespace cvlhip
