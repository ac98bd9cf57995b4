// partition.hpp -- permutation engine shared by permute and rank.
#pragma once

#include "cvl_hip.hpp"

namespace cvlhip {

// d[q(k)] = s[k] for a permutation q: q(k) = p[k], or offs[seg(k)] + p[k]
// when sd (a descriptor of n elements in m > 1 segments) is given.
template <class T>
void partition_permute(void *d, const void *s, const int *p, const void *sd, int64_t n, int64_t m, void *scratch);
int64_t partition_permute_scratch_bytes(int64_t n, int width);
// rank[order[k]] = k, or k minus the start of k's segment when sd (n
// elements, m > 1 segments) is given: the inverse of a sort order.
void partition_invert(int *rank, const int *order, int64_t n, const void *sd, int64_t m, void *scratch);

// The same inverse (unsegmented) in two steps, for a caller that places
// the sorted elements itself: partition_invert_begin lays out `scratch`
// (partition_permute_scratch_bytes(n, 4)) and clears its counters; the
// caller then writes, for every position k, the pair (order[k], k) into
// child order[k] >> clog of `pairs` -- child c's pairs at [c << clog,
// (c << clog) + cnt[c]), cnt[] the children's totals, which the caller
// also writes; partition_invert_finish runs the remaining passes.  With
// pairs == nullptr (n of at most 2^21 elements: one partition pass or
// none) the caller uses partition_invert instead.
struct InvertFirstPass {
    uint2 *pairs;
    unsigned *cnt;
    int clog;
};
InvertFirstPass partition_invert_begin(int64_t n, void *scratch);
void partition_invert_finish(int *rank, int64_t n, void *scratch);

}  // namespace cvlhip
