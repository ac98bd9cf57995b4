// segstrat.hip -- the segmented-op schedules as entry points of their own
// (for the autotuner, which times each one), and the descriptor statistics
// the dispatcher network reads.
//
// The schedules themselves are in segsched.hpp: flat look-back tiles of 8,
// 2 or 1 rows per wave, thread-, wave- or workgroup-per-segment.  The CVL
// entries (*_se*, *_re*) run the one mke_fov picked for their descriptor;
// cvlhip_seg_scan / cvlhip_seg_reduce run the arm they are given, through
// the same kernel, so a timing taken here is the time of that pick.
//
// Statistics (cvlhip_segd_features): the features the network reads, from
// the length-class histogram mke_fov stored in the descriptor's header
// (mlp.hpp, sched_feature), computed on the device by the same code the
// pick uses.
#include "cvl_hip.hpp"
#include "segd.hpp"
#include "scan.hpp"
#include "segsched.hpp"
#include "mlp.hpp"
#include "../../include/cvl.h"
#include "../../include/cvl_hip_ext.h"

namespace cvlhip {

static void dispatch(int op, int ty, bool reduce, int arm, void *d, const void *s, const void *sd, int64_t n,
                     int64_t m, void *scratch) {
    // op: 0 add, 1 mul, 2 min, 3 max; ty: 0 int, 1 double
    if (arm < 0 || arm >= kSchedArms) { usage_error("segmented schedule %d does not exist", arm); return; }
#define CVLHIP_ARM(OP)                                                         \
    do {                                                                       \
        if (reduce) reduce_seg<OP>(d, s, sd, n, m, scratch, arm);              \
        else scan_seg<OP>(d, s, sd, n, m, scratch, arm);                       \
        return;                                                                \
    } while (0)
    if (ty == 0) {
        switch (op) {
        case 0: CVLHIP_ARM(AddZ);
        case 1: CVLHIP_ARM(MulZ);
        case 2: CVLHIP_ARM(MinZ);
        case 3: CVLHIP_ARM(MaxZ);
        }
    } else if (ty == 1) {
        switch (op) {
        case 0: CVLHIP_ARM(AddD);
        case 1: CVLHIP_ARM(MulD);
        case 2: CVLHIP_ARM(MinD);
        case 3: CVLHIP_ARM(MaxD);
        }
    }
#undef CVLHIP_ARM
    usage_error("segmented op %d / type %d is not scheduled here", op, ty);
}

// Features of the descriptor `sd` (a scan of 4-byte values: the caller
// sets the last two, value size and kind).
__global__ __launch_bounds__(kWave) void k_segd_features(const SegdHeader *__restrict__ hdr, int64_t n, int64_t m,
                                                         float *__restrict__ out) {
    const int lane = threadIdx.x;
    if (lane < kF) out[lane] = sched_feature(lane, n, m, hdr->hist, hdr->hist + kLenClasses, false, false);
}

static float *g_feat_dev = nullptr;
static float *g_feat_host = nullptr;

}  // namespace cvlhip

using namespace cvlhip;

extern "C" {

int cvlhip_seg_arms(void) { return kSchedArms; }
int cvlhip_seg_features_count(void) { return kF; }

void cvlhip_seg_scan(int op, int ty, int arm, vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch) {
    dispatch(op, ty, false, arm, d, s, sd, n, m, scratch);
}

void cvlhip_seg_reduce(int op, int ty, int arm, vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch) {
    dispatch(op, ty, true, arm, d, s, sd, n, m, scratch);
}

int cvlhip_seg_scratch(int n, int m) {
    (void)m;
    int64_t a = scan_scratch_bytes<AddD>(n);
    const int64_t c[] = {sched_scratch_bytes<AddZ, false>(n), sched_scratch_bytes<AddD, false>(n),
                         sched_scratch_bytes<AddZ, true>(n), sched_scratch_bytes<AddD, true>(n),
                         kReduceScratchBytes};
    for (int64_t v : c) a = v > a ? v : a;
    return units_for_bytes(a);
}

// Features of a descriptor built by mke_fov, written to host memory `out`
// (kF floats, see mlp.hpp); [45] and [46] (8-byte values, reduction) are 0
// here and set by the caller.
void cvlhip_segd_features(vec_p sd, int n, int m, float *out) {
    if (!g_feat_dev) {
        HIPCHECK(hipMalloc(&g_feat_dev, sizeof(float) * kF));
        HIPCHECK(hipHostMalloc(&g_feat_host, sizeof(float) * kF, hipHostMallocDefault));
    }
    k_segd_features<<<1, kWave, 0, g_stream>>>((const SegdHeader *)sd, n, m, g_feat_dev);
    LAUNCHCHECK();
    HIPCHECK(hipMemcpyAsync(g_feat_host, g_feat_dev, sizeof(float) * kF, hipMemcpyDeviceToHost, g_stream));
    HIPCHECK(hipStreamSynchronize(g_stream));
    for (int i = 0; i < kF; i++) out[i] = g_feat_host[i];
}

// The arm the descriptor's header names for a scan (reduce = 0) or a
// reduction of 4- or 8-byte values, or -1 when it names none.
int cvlhip_segd_pick(vec_p sd, int reduce, int elem_bytes) {
    SegdHeader h;
    HIPCHECK(hipMemcpyAsync(&h, sd, sizeof h, hipMemcpyDeviceToHost, g_stream));
    HIPCHECK(hipStreamSynchronize(g_stream));
    if ((h.sched >> 16) != kSchedMagic) return -1;
    return (int)((h.sched >> (4 * sched_slot(reduce != 0, elem_bytes))) & 15u);
}

}  // extern "C"
