// Batched launch building blocks shared by Rav2dGpuBatch and the server.
#pragma once
#include "frame.h"

namespace r2g {

constexpr int kPasses = RAV2D_NUM_PASSES;
// bit-depth groups: 8-, 10- and 12-bit frames
constexpr int kDepths = 3;
inline int depth_group(int bitdepth) { return (bitdepth - 8) >> 1; }
inline int group_depth(int g) { return 8 + 2 * g; }

// Host-side launch arguments of a set of frames, per pass and bit-depth
// group (8-, 10-, 12-bit).
struct HostBatch {
    // deblock_rows > 0: both deblocking passes of a frame run as one tile pass
    // (DeblockVHArgs, tiles of 64 x deblock_rows) in the vertical-edge pass's
    // slot, and the horizontal-edge pass has no items
    int deblock_rows = 0;
    bool any_ss = false;  // a frame has 4:2:2 / 4:4:4 chroma: the general deblock / CDEF / grain kernels
    std::vector<DeblockArgs> dbv[kDepths], dbh[kDepths];
    std::vector<DeblockVHArgs> dbvh[kDepths];
    std::vector<CdefArgs> cdef[kDepths];
    std::vector<LrArgs> lr[kDepths];
    std::vector<GrainArgs> gapply[kDepths], gsetup;
    std::vector<int32_t> prefix[kPasses][kDepths];
    int64_t total[kPasses][kDepths] = {};  // work items of each launch
    int64_t bytes[kPasses] = {};     // algorithmic bytes: one read + one write of the planes
    int64_t pixels = 0;              // luma pixels of the frames
    int frames[kPasses] = {};        // frames in each pass
    int64_t items(int p) const {
        int64_t n = 0;
        for (int g = 0; g < kDepths; g++) n += total[p][g];
        return n;
    }
};

// Device view of a packed HostBatch.
struct DevBatch {
    int deblock_rows = 0;
    bool any_ss = false;  // a frame has 4:2:2 / 4:4:4 chroma: the general deblock / CDEF / grain kernels
    DeblockArgs *dbv[kDepths] = {}, *dbh[kDepths] = {};
    DeblockVHArgs *dbvh[kDepths] = {};
    CdefArgs *cdef[kDepths] = {};
    LrArgs *lr[kDepths] = {};
    GrainArgs *gapply[kDepths] = {}, *gsetup = nullptr;
    int32_t *prefix[kPasses][kDepths] = {};
    int nf[kPasses][kDepths] = {};
    int32_t total[kPasses][kDepths] = {};
    int n_setup = 0;
    int64_t items(int p) const {
        int64_t n = 0;
        for (int g = 0; g < kDepths; g++) n += total[p][g];
        return n;
    }
};

// Launch shape of each pass: workgroups (0 = one per item) and XCD remap;
// mixed_depths: one launch per pass over every bit-depth group (else one per
// group present).
struct LaunchShape {
    int32_t grid[kPasses];
    int32_t xcd_remap[kPasses];
    int32_t mixed_depths = 0;
};

// Plans the frames' chains (deblock, CDEF, LR into OUT; with apply_grain,
// grain into DISP) and gathers their launch arguments; deblock_rows 32 or 64
// fuses the two deblocking passes (0: separate launches).
int plan_batch(Rav2dGpuCtx *ctx, Rav2dGpuFrame *const *frames, int n, bool apply_grain, HostBatch &hb,
               int deblock_rows = 0);
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
    // high-bit-depth groups' (untimed, unprofiled runs only)
    hipStream_t side2 = nullptr;
    hipEvent_t ev_fork2 = nullptr, ev_join2 = nullptr;
    int time_pass = -1;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    Rav2dGpuCtx *profile = nullptr;  // context whose profiling events bracket each pass
};

int enqueue_batch(const DevBatch &db, const LaunchShape &shape, bool apply_grain, const EnqueueCtx &ec);

}  // namespace r2g
