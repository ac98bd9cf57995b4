// scan.hpp -- scan/reduce entry points shared between translation units.
#pragma once

#include "cvl_hip.hpp"

namespace cvlhip {

template <class Op> void scan_unseg(void *d, const void *s, int64_t n, void *scratch, unsigned *err = nullptr);
// Segmented scan / reduction under the descriptor's schedule, or under arm
// `arm` (0 .. kSchedArms-1, segsched.hpp) when arm >= 0.
template <class Op>
void scan_seg(void *d, const void *s, const void *sd, int64_t n, int64_t m, void *scratch, int arm = -1,
              unsigned *err = nullptr);
template <class Op>
void reduce_seg(void *d, const void *s, const void *sd, int64_t n, int64_t m, void *scratch, int arm = -1);
template <class Op> typename Op::T reduce_unseg(const void *s, int64_t n, void *scratch);
template <class Op> int64_t scan_scratch_bytes(int64_t n);
// Scratch of a segmented reduction of n elements of elem_bytes each, under
// any schedule.
int64_t seg_reduce_scratch_bytes(int64_t n, int elem_bytes);

// Scratch of an unsegmented reduction: a 256-byte header, one 8-byte slot
// per block partial, and the result slot (at kReduceResultOffset).
constexpr int kReducePartials = 2048;     // most blocks of one reduction
constexpr int64_t kReduceResultOffset = 256 + (int64_t)kReducePartials * 8;
constexpr int64_t kReduceScratchBytes = kReduceResultOffset + 256;

}  // namespace cvlhip
