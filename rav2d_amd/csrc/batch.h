// Batched launch building blocks shared by Rav2dGpuBatch and the server.
#pragma once
#include "frame.h"

namespace r2g {

constexpr int kPasses = RAV2D_NUM_PASSES;

// Host-side launch arguments of a set of frames.  Each pass has one launch
// per bit depth present (group 0: 8-bit frames, group 1: 10-bit frames), so
// every launch runs one sample-type instantiation of the kernel body.
struct HostBatch {
    std::vector<DeblockArgs> dbv[2], dbh[2];
    std::vector<CdefArgs> cdef[2];
    std::vector<LrArgs> lr[2];
    std::vector<GrainArgs> gapply[2], gsetup;
    std::vector<int32_t> prefix[kPasses][2];
    int64_t total[kPasses][2] = {};  // work items of each launch
    int64_t bytes[kPasses] = {};     // algorithmic bytes: one read + one write of the planes
    int64_t pixels = 0;              // luma pixels of the frames
    int frames[kPasses] = {};        // frames in each pass
    int64_t items(int p) const { return total[p][0] + total[p][1]; }
};

// Device view of a packed HostBatch.
struct DevBatch {
    DeblockArgs *dbv[2] = {}, *dbh[2] = {};
    CdefArgs *cdef[2] = {};
    LrArgs *lr[2] = {};
    GrainArgs *gapply[2] = {}, *gsetup = nullptr;
    int32_t *prefix[kPasses][2] = {};
    int nf[kPasses][2] = {};
    int32_t total[kPasses][2] = {};
    int n_setup = 0;
};

// Launch shape of each pass: workgroups (0 = one per item) and XCD remap.
struct LaunchShape {
    int32_t grid[kPasses];
    int32_t xcd_remap[kPasses];
};

// Plans the frames' chains (deblock, CDEF, LR into OUT; with apply_grain,
// grain into DISP) and gathers their launch arguments.
int plan_batch(Rav2dGpuCtx *ctx, Rav2dGpuFrame *const *frames, int n, bool apply_grain, HostBatch &hb);
// Bytes of the packed argument block.
size_t packed_size(const HostBatch &hb);
// An upper bound of packed_size for any batch of at most n frames.
size_t packed_bound(int n);
// Packs hb into `host` (packed_size bytes) laid out for device address `dev`.
void pack_batch(const HostBatch &hb, uint8_t *host, uint8_t *dev, DevBatch &db);
// Makes the constant tables of the current device ready (outside any capture).
int batch_init_device();

// Stream resources of one enqueue: `side` runs the grain set-up, forked from
// and joined back into `s` with ev_fork / ev_join.  ev_a / ev_b, when set,
// bracket the launch of pass `time_pass`.
struct EnqueueCtx {
    hipStream_t s, side;
    hipEvent_t ev_fork, ev_join;
    // optional: a second stream that runs the 8-bit group's passes beside the
    // 10-bit group's (untimed, unprofiled runs only)
    hipStream_t side2 = nullptr;
    hipEvent_t ev_fork2 = nullptr, ev_join2 = nullptr;
    int time_pass = -1;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    Rav2dGpuCtx *profile = nullptr;  // context whose profiling events bracket each pass
};

int enqueue_batch(const DevBatch &db, const LaunchShape &shape, bool apply_grain, const EnqueueCtx &ec);

}  // namespace r2g
