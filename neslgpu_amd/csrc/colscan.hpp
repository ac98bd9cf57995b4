// colscan.hpp -- exclusive scan of each row of a digits x tiles count
// matrix (row-major, `rows` rows of `cols` counts), used by the chain-free
// radix passes: after it, entry [d][t] is the number of elements with digit
// d in tiles before t.  One block per row; the row total goes to totals[d]
// when totals is non-null.
#pragma once

#include "cvl_hip.hpp"

namespace cvlhip {

constexpr int kColScanThreads = 1024;

__global__ __launch_bounds__(kColScanThreads) inline void k_col_scan(unsigned *__restrict__ m, int64_t cols,
                                                                     unsigned *__restrict__ totals) {
    __shared__ unsigned s_w[kColScanThreads / kWave];
    unsigned *row = m + (int64_t)blockIdx.x * cols;
    const int tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
    unsigned carry = 0;
    for (int64_t c0 = 0; c0 < cols; c0 += kColScanThreads) {
        const int64_t c = c0 + tid;
        const unsigned v = c < cols ? row[c] : 0u;
        const unsigned inc = wave_inclusive_scan(v, AddU());
        if (lane == kWave - 1) s_w[w] = inc;
        __syncthreads();
        unsigned before = 0, all = 0;
        for (int i = 0; i < kColScanThreads / kWave; i++) {
            const unsigned x = s_w[i];
            if (i < w) before += x;
            all += x;
        }
        if (c < cols) row[c] = carry + before + inc - v;
        carry += all;
        __syncthreads();
    }
    if (totals && tid == 0) totals[blockIdx.x] = carry;
}

}  // namespace cvlhip
