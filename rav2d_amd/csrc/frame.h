// Host-side frame and context objects shared by the runtime sources.
#pragma once
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "internal.h"

namespace r2g {

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            set_last_error(std::string(#expr) + ": " + hipGetErrorString(e_));                 \
            return RAV2D_EDEVICE;                                                              \
        }                                                                                      \
    } while (0)


}  // namespace r2g

struct Rav2dGpuCtx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;  // parameter-only work (grain tables) overlapped with the passes
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t budget = 0;
    uint64_t in_use = 0;
    std::mutex mu;
    // per-pass profiling: event pairs recorded around launches, resolved lazily
    bool profiling = false;
    struct Rec {
        int pass;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> free_events;
    double pass_ms[RAV2D_NUM_PASSES] = {0};
    int64_t pass_count[RAV2D_NUM_PASSES] = {0};
    // device buffers handed out by rav2d_gpu_alloc / rav2d_gpu_picture_alloc
    std::unordered_map<void *, size_t> allocs;
};

namespace r2g {

inline hipEvent_t take_event(Rav2dGpuCtx *c) {
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Events that only order work between streams of one device: no timing and no
// system-scope fence (which would write back and invalidate the L2 caches at
// every fork and join); the kernels' own device-scope fences order the data.
constexpr unsigned kStreamSyncEventFlags = hipEventDisableTiming | hipEventDisableSystemFence;

// Times one pass launch when profiling is on: arm() hands two events to the
// next kernel launched on this thread (its own start and end, launch gaps
// excluded); the constructor arms unless told not to (a pass whose timed
// kernel follows another launch arms just before it).
struct PassTimer {
    Rav2dGpuCtx *c;
    int pass;
    hipEvent_t a = nullptr, b = nullptr;
    PassTimer(Rav2dGpuCtx *c_, int pass_, hipStream_t s, bool arm_now = true) : c(c_), pass(pass_) {
        if (c && c->profiling && s == c->stream) {
            a = take_event(c);
            b = take_event(c);
            if (arm_now) arm();
        }
    }
    void arm() {
        if (!a || !b) return;
        LaunchTiming &lt = launch_timing();
        lt.start = a;
        lt.stop = b;
        lt.used = false;
    }
    ~PassTimer() {
        LaunchTiming &lt = launch_timing();
        const bool used = a && b && lt.used && !lt.start;
        lt.start = lt.stop = nullptr;
        lt.used = false;
        if (used) {
            c->recs.push_back({pass, a, b});
        } else {
            if (a) c->free_events.push_back(a);
            if (b) c->free_events.push_back(b);
        }
    }
};

}  // namespace r2g

// Buffer roles of a frame.  F0 is never written by a pass.
enum { BUF_IN = 0, BUF_OUT = 1, BUF_DISP = 2, BUF_T1 = 3, BUF_T2 = 4, NBUF = 5 };

struct Rav2dGpuFrame {
    Rav2dGpuCtx *ctx = nullptr;
    Rav2dFrameParams params;
    r2g::DevGeom g;
    int pw[3], ph[3], bps;
    int64_t stride[3];            // bytes, per plane
    size_t plane_bytes[3];
    uint8_t *block = nullptr;     // one allocation: all plane buffers
    uint8_t *buf[NBUF][3];
    // caller-owned planes standing in for buf[b] during one call
    // (rav2d_gpu_filter_planes); null = the frame's own
    uint8_t *ext[NBUF][3] = {};
    Rav2dPackedFilter *d_mask = nullptr;
    r2g::ChromaMasks *d_uvm = nullptr;  // 4:2:2 / 4:4:4 frames: full-size chroma masks per SB256
    size_t n_mask = 0;
    size_t n_lr = 0;
    r2g::DbLut *d_dblut = nullptr;
    // segment maps (rav2d_gpu_frame_set_segmap): luma ids bw x bh, then chroma
    // ids (bw >> ss_hor) x (bh >> ss_ver), packed; allocated at the first set
    uint8_t *d_seg = nullptr;
    bool seg_ok = false;
    size_t bytes = 0;
    bool side_ok = false;
    // loop-restoration work list (rebuilt when params or side info change)
    std::vector<Rav2dPackedFilter> h_mask;  // host copy (after the bottom crop)
    std::vector<r2g::ChromaMasks> h_uvm;    // host copy of d_uvm (4:2:2 / 4:4:4)
    std::vector<Rav2dAv2Restoration> h_lr;
    std::vector<r2g::LrItem> lr_items;
    r2g::LrItem *d_lr_items = nullptr;
    size_t d_lr_cap = 0;
    int32_t *d_cdef_order = nullptr;  // launch order of the 64x64 CDEF superblocks
    int lr_restore_planes = 0, lr_base_q = 0, lr_widx = 0, lr_ns_idx = 0;
    int lr_n_restore = 0;  // lr_items: restoration items first, then plane copies
    bool lr_dirty = true;
    uint8_t *d_grain = nullptr;  // grain LUTs + scaling + offsets scratch
    uint32_t *d_grain_err = nullptr;  // error flags of the grain set-up (last word of d_grain)
    hipEvent_t ev_setup = nullptr, ev_free = nullptr;
    // recorded after the frame's last per-frame launch (filter, grain, pass):
    // parameter and side-info changes wait for it before rewriting device tables
    hipEvent_t ev_done = nullptr;
    bool done_pending = false;
    bool grain_ready = false, grain_used = false;
    uint64_t gen = 0;  // bumped when parameters or side info change
};

namespace r2g {

// Per-pass launch arguments of one frame's in-loop chain.
struct ChainPlan {
    bool on[4];         // deblock V, deblock H, CDEF, LR launched for this frame
    DeblockArgs dbv, dbh;
    CdefArgs cdef;
    LrArgs lr;
    int n_lr;           // items of the LR launch (restoration units and plane copies)
    int final_buf[3];   // buffer that holds each plane after the launched passes
};

int plan_chain(Rav2dGpuCtx *ctx, Rav2dGpuFrame *f, int upto, ChainPlan &cp);
DevPlane plane_of(const Rav2dGpuFrame *f, int b, int pl);
// Planes of a picture of the given layout: 1 for monochrome (I400), else 3.
inline int n_planes(int layout) { return layout == RAV2D_PIXEL_LAYOUT_I400 ? 1 : 3; }
// Samples of picture padding right of each plane of a width-w picture: blocks
// can overhang the right edge only when w is not a multiple of 128 (the
// largest superblock).
inline int pad_columns(int w) { return (w & 127) ? RAV2D_PICTURE_PAD : 0; }
bool frame_has_grain(const Rav2dFrameParams &P);
int grain_args(Rav2dGpuFrame *f, GrainArgs &a);
bool data_has_grain(const Rav2dFilmGrainData &d);
// Bytes of a grain scratch area (templates, scaling LUTs, block offsets, error
// word) for a width x height picture, and the arguments of the grain kernels
// for one with its scratch at `scratch` (src / dst planes left to the caller).
size_t grain_scratch_bytes(int width, int height);
int grain_args_for(const Rav2dFilmGrainData &d, uint32_t seed, int bitdepth, int width, int height, uint8_t *scratch,
                   uint32_t *err, GrainArgs &a);
int frame_errors(Rav2dGpuFrame *f);  // after a synchronization: RAV2D_EDEVICE if a kernel flagged an error
hipStream_t pick_stream(Rav2dGpuCtx *ctx, void *stream);

}  // namespace r2g
