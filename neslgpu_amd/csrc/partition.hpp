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

}  // namespace cvlhip
