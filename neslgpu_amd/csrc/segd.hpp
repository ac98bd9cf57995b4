// segd.hpp -- the segment descriptor of the HIP backend.
//
// The serial CVL's descriptor is the vector of m segment lengths
// (ref nesl/cvl/serial/facilt.c:106-123).  Here the descriptor holds a
// small header with its own shape, those lengths, and data derived once by
// mke_fov, which the segmented kernels use to find segment boundaries in
// O(1):
//
//   hdr              SegdHeader: int64 n, int64 m (a kernel handed only the
//                    descriptor, such as a generated fused operation, reads
//                    its shape), the length-class histogram and the
//                    schedules the dispatcher picked for it
//   lens[m]          segment lengths (int32)
//   offs[m+1]        exclusive prefix of lens; offs[m] = n
//   nz[m]            indices of the non-empty segments, in order
//   wcnt[W+1]        number of segment heads before each bitmap word
//   heads[W]         bit j set when a non-empty segment starts at element j
//                    (W = ceil(n/32) words, plus padding)
//
// The segment of element k is nz[rank(k) - 1] where rank(k) counts the
// heads at or before k: wcnt[k>>5] + popcount(heads[k>>5] & low bits).
// Every section starts on a 256-byte boundary.
#pragma once

#include <stdint.h>

namespace cvlhip {

// Segment-length classes of the descriptor statistics: class c holds the
// lengths in [2^c, 2^(c+1)); the last class is open above.
constexpr int kLenClasses = 21;

// The first 256 bytes of a descriptor.  mke_fov fills the histogram while
// it scans the lengths and, for descriptors of at least kSchedMinN
// elements, lets the dispatcher network pick a schedule for each kind of
// segmented operation (scan or reduction, 4- or 8-byte values); the picks
// are 4-bit arm numbers under a magic half-word, so a descriptor that was
// not built that way reads as "no pick".
struct SegdHeader {
    int64_t n, m;
    uint32_t sched;                  // kSchedMagic << 16 | picks, 4 bits each at 4 * sched_slot
    uint32_t stats;                  // kStatsMagic when the histogram below is filled
    // hist[c]: elements in segments of length class c; hist[kLenClasses + c]:
    // segments of class c (empty segments are not counted)
    uint32_t hist[2 * kLenClasses];
};
static_assert(sizeof(SegdHeader) <= 256, "the header section is 256 bytes");
constexpr uint32_t kSchedMagic = 0x5c4du;
constexpr uint32_t kStatsMagic = 0x57a75u;
// Descriptors of fewer elements get no pick: every schedule then runs in
// one or two tiles and the fixed one is used.
constexpr int64_t kSchedMinN = 1 << 15;
// Where a kind of operation finds its pick: scan/reduction x element size.
__host__ __device__ constexpr int sched_slot(bool reduce, int elem_bytes) {
    return (reduce ? 2 : 0) + (elem_bytes == 8 ? 1 : 0);
}

struct SegdLayout {
    int64_t hdr, lens, offs, nz, wcnt, heads, bytes;   // byte offsets and total size
    int64_t words;                                 // bitmap words

    __host__ __device__ static int64_t up256(int64_t b) { return (b + 255) & ~(int64_t)255; }

    __host__ __device__ SegdLayout(int64_t n, int64_t m) {
        words = (n + 31) / 32;
        hdr = 0;
        lens = 256;
        offs = up256(lens + 4 * m);
        nz = up256(offs + 4 * (m + 1));
        wcnt = up256(nz + 4 * m);
        heads = up256(wcnt + 4 * (words + 1));
        bytes = up256(heads + 4 * (words + 8));
    }
};

struct SegdView {
    const int *lens;
    const int *offs;
    const int *nz;
    const int *wcnt;
    const uint32_t *heads;
    int64_t n, m;

    __host__ __device__ SegdView()
        : lens(nullptr), offs(nullptr), nz(nullptr), wcnt(nullptr), heads(nullptr), n(0), m(0) {}
    __host__ __device__ SegdView(const void *sd, int64_t n_, int64_t m_) : n(n_), m(m_) {
        SegdLayout L(n_, m_);
        const char *b = (const char *)sd;
        lens = (const int *)(b + L.lens);
        offs = (const int *)(b + L.offs);
        nz = (const int *)(b + L.nz);
        wcnt = (const int *)(b + L.wcnt);
        heads = (const uint32_t *)(b + L.heads);
    }

    // The view of a descriptor whose shape only the descriptor knows.
    __device__ static SegdView from_header(const void *sd) {
        const int64_t *h = (const int64_t *)sd;
        return SegdView(sd, h[0], h[1]);
    }

    // Is element k the first element of a (non-empty) segment?
    __device__ __forceinline__ bool is_head(int64_t k) const {
        return (heads[k >> 5] >> (k & 31)) & 1u;
    }
    // Segment index of element k (0 <= k < n).
    __device__ __forceinline__ int seg_of(int64_t k) const {
        uint32_t w = heads[k >> 5];
        uint32_t mask = (uint32_t)(((uint64_t)2 << (k & 31)) - 1);
        int r = wcnt[k >> 5] + __popc(w & mask);
        return nz[r - 1];
    }
    // Does every segment hold an element?  Then nz is the identity.
    __device__ __forceinline__ bool all_nonempty() const { return wcnt[(n + 31) >> 5] == m; }
    // seg_of, skipping the nz lookup when all_nonempty() holds.
    __device__ __forceinline__ int seg_of(int64_t k, bool nonempty) const {
        uint32_t w = heads[k >> 5];
        uint32_t mask = (uint32_t)(((uint64_t)2 << (k & 31)) - 1);
        int r = wcnt[k >> 5] + __popc(w & mask);
        return nonempty ? r - 1 : nz[r - 1];
    }
};

}  // namespace cvlhip
