// tuned.hpp -- tile and workgroup shapes chosen by the autotuner
// (python -m neslgpu_amd.tune pick), force-included into every HIP
// source by neslgpu_amd/build.py.  Each value is the fastest correct
// one of its macro's candidates on MI355X, the others at their values.
#pragma once
#ifndef CVLHIP_SCAN_ROWS_U4
#define CVLHIP_SCAN_ROWS_U4 12
#endif
#ifndef CVLHIP_SCAN_BLOCK_U4
#define CVLHIP_SCAN_BLOCK_U4 1024
#endif
#ifndef CVLHIP_LOOKBACK_K
#define CVLHIP_LOOKBACK_K 1
#endif
#ifndef CVLHIP_SEG_SCAN_BLOCK4
#define CVLHIP_SEG_SCAN_BLOCK4 512
#endif
#ifndef CVLHIP_SEG_SCAN_BLOCK8
#define CVLHIP_SEG_SCAN_BLOCK8 1024
#endif
#ifndef CVLHIP_REDUCE_GRID
#define CVLHIP_REDUCE_GRID 2048
#endif
#ifndef CVLHIP_PACK_ROWS
#define CVLHIP_PACK_ROWS 8
#endif
#ifndef CVLHIP_PART_ITEMS
#define CVLHIP_PART_ITEMS 8
#endif
#ifndef CVLHIP_PART_THREADS
#define CVLHIP_PART_THREADS 1024
#endif
#ifndef CVLHIP_BLK_LOG4
#define CVLHIP_BLK_LOG4 13
#endif
#ifndef CVLHIP_PART_ATOMIC_FIRST
#define CVLHIP_PART_ATOMIC_FIRST 0
#endif
#ifndef CVLHIP_RANK_THREADS
#define CVLHIP_RANK_THREADS 512
#endif
#ifndef CVLHIP_RANK_ITEMS
#define CVLHIP_RANK_ITEMS 8
#endif
#ifndef CVLHIP_RANK_CHAIN
#define CVLHIP_RANK_CHAIN 0
#endif
#ifndef CVLHIP_TILE_SCAN_ROWS
#define CVLHIP_TILE_SCAN_ROWS 64
#endif
#ifndef CVLHIP_SEG_REDUCE_ROWS
#define CVLHIP_SEG_REDUCE_ROWS 2
#endif
#ifndef CVLHIP_SEG_REDUCE_BLOCK
#define CVLHIP_SEG_REDUCE_BLOCK 256
#endif
#ifndef CVLHIP_SEG_REDUCE_LANE_BYTES
#define CVLHIP_SEG_REDUCE_LANE_BYTES 64
#endif
#ifndef CVLHIP_PAIR_ROWS
#define CVLHIP_PAIR_ROWS 8
#endif
#ifndef CVLHIP_PAIR_BLOCK
#define CVLHIP_PAIR_BLOCK 512
#endif
#ifndef CVLHIP_PAIR_LANE_BYTES
#define CVLHIP_PAIR_LANE_BYTES 16
#endif
