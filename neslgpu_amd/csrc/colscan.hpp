// colscan.hpp -- exclusive scan down the columns of a tiles x 256 count
// matrix, used by the chain-free radix passes.  Row t holds tile t's count
// of each of the 256 digits (tile-major: a histogram kernel writes its row as
// one contiguous 1 KiB store, and the reordering kernel reads it back the
// same way).  After the scan, entry [t][d] is the number of elements with
// digit d in tiles before t, totals[d] (when non-null) the column total and
// starts[d] (when non-null) the sum of the totals of the columns before d.
//
// One kernel: workgroup g owns rows [g*R, (g+1)*R) and thread d column d.
// It sums its R counts, publishes the sum in a status word (flag | value,
// one 64-bit word per (group, digit)), walks back over earlier groups' words
// until one holds an inclusive prefix, publishes its own inclusive prefix and
// writes the R exclusive prefixes.  Every group publishes before it waits
// and groups start in order, so every walk ends; the spin limit only guards
// against a corrupted status array.
#pragma once

#include "cvl_hip.hpp"
#include "lookback.hpp"

namespace cvlhip {

#ifndef CVLHIP_TILE_SCAN_ROWS
#define CVLHIP_TILE_SCAN_ROWS 64
#endif
constexpr int kTileScanRows = CVLHIP_TILE_SCAN_ROWS;   // rows (tiles) per workgroup
constexpr int kTileScanCols = 256;

__global__ __launch_bounds__(kTileScanCols) inline void k_tile_scan(unsigned *__restrict__ m, int64_t tiles,
                                                                    unsigned *__restrict__ totals,
                                                                    uint64_t *__restrict__ status,
                                                                    unsigned *__restrict__ err,
                                                                    unsigned *__restrict__ starts) {
    constexpr int R = kTileScanRows;
    const int d = threadIdx.x;
    const int64_t g = blockIdx.x;
    const int64_t r0 = g * R;
    unsigned v[R];
    unsigned sum = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        v[r] = r0 + r < tiles ? m[(r0 + r) * kTileScanCols + d] : 0u;
        sum += v[r];
    }
    uint64_t *mine = status + g * kTileScanCols + d;
    unsigned excl = 0;
    if (g == 0) {
        st_status(mine, ((uint64_t)kPrefix << 32) | sum);
    } else {
        st_status(mine, ((uint64_t)kAggregate << 32) | sum);
        // Walk back over the earlier groups, LB status words loaded at once
        // (independent loads: one latency per LB groups) and taken in order
        // up to the first prefix; a group not yet published stops the batch
        // and the walk resumes there.  Group 0 always publishes a prefix,
        // so the walk ends at t = 0 at the latest.
        constexpr int LB = 8;
        int64_t t = g - 1;
        int spins = 0;
        for (;;) {
            uint64_t w[LB];
#pragma unroll
            for (int k = 0; k < LB; k++) w[k] = t - k >= 0 ? ld_status(status + (t - k) * kTileScanCols + d) : 0;
            int used = 0;          // words taken from this batch
            bool done = false, wait = false;
#pragma unroll
            for (int k = 0; k < LB; k++) {
                if (done || wait) continue;
                const uint32_t f = (uint32_t)(w[k] >> 32);
                if (f == kInvalid) {
                    wait = true;
                    continue;
                }
                excl += (uint32_t)w[k];
                used = k + 1;
                if (f == kPrefix) done = true;
            }
            if (done) break;
            t -= used;
            if (used > 0) spins = 0;
            if (wait) {
                if (++spins >= kSpinLimit) {
                    atomicOr(err, 2u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        st_status(mine, ((uint64_t)kPrefix << 32) | (excl + sum));
    }
    unsigned run = excl;
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (r0 + r < tiles) m[(r0 + r) * kTileScanCols + d] = run;
        run += v[r];
    }
    if (g == (int64_t)gridDim.x - 1) {
        if (totals) totals[d] = run;
        if (starts) {
            // the columns' starts: exclusive scan of the totals across them
            __shared__ unsigned s_w[kTileScanCols / kWave];
            const int w = d / kWave, lane = d % kWave;
            const unsigned inc = wave_inclusive_scan(run, AddU());
            if (lane == kWave - 1) s_w[w] = inc;
            __syncthreads();
            unsigned base = 0;
            for (int i = 0; i < w; i++) base += s_w[i];
            starts[d] = base + inc - run;
        }
    }
}

// Status words a scan of `tiles` rows needs (64-bit words).
inline int64_t tile_scan_status_words(int64_t tiles) {
    return (int64_t)kTileScanCols * ((tiles + kTileScanRows - 1) / kTileScanRows + 1);
}

// Scan m (tiles x 256) in place on the library stream; `status` holds
// tile_scan_status_words(tiles) words and is cleared here unless the caller
// cleared it already (`cleared`).  `starts` (optional) receives the
// exclusive scan of the column totals: where each column's run begins.
inline void tile_scan(unsigned *m, int64_t tiles, unsigned *totals, uint64_t *status, bool cleared = false,
                      unsigned *starts = nullptr) {
    if (tiles <= 0) return;
    const int64_t groups = (tiles + kTileScanRows - 1) / kTileScanRows;
    if (!cleared) fill_u32(status, 0, 2 * kTileScanCols * groups);
    k_tile_scan<<<(unsigned)groups, kTileScanCols, 0, g_stream>>>(m, tiles, totals, status, device_error_word(),
                                                                  starts);
    LAUNCHCHECK();
}

}  // namespace cvlhip
