// scan.hpp -- scan/reduce entry points shared between translation units.
#pragma once

#include "cvl_hip.hpp"

namespace cvlhip {

template <class Op> void scan_unseg(void *d, const void *s, int64_t n, void *scratch);
template <class Op> void scan_seg(void *d, const void *s, const void *sd, int64_t n, int64_t m, void *scratch);
template <class Op> void scan_seg_r8(void *d, const void *s, const void *sd, int64_t n, int64_t m, void *scratch);
template <class Op> void reduce_seg_r8(void *d, const void *s, const void *sd, int64_t n, int64_t m, void *scratch);
template <class Op> void reduce_seg(void *d, const void *s, const void *sd, int64_t n, int64_t m, void *scratch);
template <class Op> typename Op::T reduce_unseg(const void *s, int64_t n, void *scratch);
template <class Op> int64_t scan_scratch_bytes(int64_t n);

}  // namespace cvlhip
