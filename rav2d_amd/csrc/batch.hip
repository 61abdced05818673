// Batched filtering: many frames of any size and bit depth, one launch per
// pass and bit depth over all of them, optionally captured once and replayed
// as a hipGraph.
//
// Each launch gets an array of per-frame argument blocks and the prefix sum of
// the frames' work items (tiles, superblocks, restoration units); a workgroup
// finds its frame by binary search over the prefix.  8-bit and 10-bit frames
// go to separate launches of the pass (each instantiates one sample type, so
// its register budget is the per-frame kernel's).  The launch shape (one workgroup per item,
// or a persistent grid that loops; XCD-contiguous item ranges or round robin)
// is per pass and is what rav2d_gpu_batch_autotune picks.
//
// Pass order per frame is that of the per-frame chain (plan_chain): deblock V,
// deblock H, CDEF/CCSO, loop restoration into the output planes, film grain
// into the display planes.  LR's launch also carries copy items for every
// plane it does not restore (part of each frame's item list), so one launch
// leaves every frame's output planes complete; grain's launch copies the
// planes of frames without grain.  The
// grain tables of all grained frames are built by one set-up launch on a side
// stream that forks at the start of the batch and joins before grain.
#include <string.h>

#include <algorithm>
#include <vector>

#include "batch.h"

namespace r2g {

namespace {

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct Layout {
    size_t dbv[kDepths], dbh[kDepths], dbvh[kDepths], cdef[kDepths], lr[kDepths], gapply[kDepths], gsetup, prefix[kPasses][kDepths],
        total;
};

Layout layout_of(const HostBatch &hb) {
    Layout L;
    size_t off = 0;
    auto reserve = [&](size_t bytes) {
        size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    for (int g = 0; g < kDepths; g++) {
        L.dbv[g] = reserve(hb.dbv[g].size() * sizeof(DeblockArgs));
        L.dbh[g] = reserve(hb.dbh[g].size() * sizeof(DeblockArgs));
        L.dbvh[g] = reserve(hb.dbvh[g].size() * sizeof(DeblockVHArgs));
        L.cdef[g] = reserve(hb.cdef[g].size() * sizeof(CdefArgs));
        L.lr[g] = reserve(hb.lr[g].size() * sizeof(LrArgs));
        L.gapply[g] = reserve(hb.gapply[g].size() * sizeof(GrainArgs));
        for (int p = 0; p < kPasses; p++) L.prefix[p][g] = reserve(hb.prefix[p][g].size() * sizeof(int32_t));
    }
    L.gsetup = reserve(hb.gsetup.size() * sizeof(GrainArgs));
    L.total = std::max<size_t>(off, 256);
    return L;
}

int grid_of(int64_t total, const LaunchShape &sh, int pass) {
    int g = sh.grid[pass];
    int64_t n = g > 0 ? std::min<int64_t>(g, total) : total;
    if (sh.xcd_remap[pass]) n = (n + 7) & ~int64_t(7);
    return (int)std::max<int64_t>(n, 1);
}

}  // namespace

int batch_init_device() {
    if (!grain_init_tables()) {
        set_last_error("film grain tables: device allocation failed");
        return RAV2D_ENOMEM;
    }
    return RAV2D_OK;
}

int plan_batch(Rav2dGpuCtx *ctx, Rav2dGpuFrame *const *frames, int n, bool apply_grain, HostBatch &hb,
               int deblock_rows) {
    hb = HostBatch();
    if (deblock_rows && !deblock_vh_rows_ok(deblock_rows)) {
        set_last_error("fused deblocking takes tiles of 32 or 64 rows");
        return RAV2D_EINVAL;
    }
    hb.deblock_rows = deblock_rows;
    for (int i = 0; i < n; i++) {
        Rav2dGpuFrame *f = frames[i];
        ChainPlan cp;
        int r = plan_chain(ctx, f, RAV2D_PASS_LR, cp);
        if (r) return r;
        const int g = depth_group(f->g.bitdepth);
        hb.any_ss = hb.any_ss || f->d_uvm != nullptr;
        // one read and one write of the frame's 1.5 * W * H samples
        const int64_t fbytes = 3 * (int64_t)f->pw[0] * f->ph[0] * f->bps;
        hb.pixels += (int64_t)f->pw[0] * f->ph[0];
        auto add = [&](int pass, int64_t count) {
            hb.prefix[pass][g].push_back((int32_t)hb.total[pass][g]);
            hb.total[pass][g] += count;
            hb.bytes[pass] += fbytes;
            hb.frames[pass]++;
        };
        if (deblock_rows && (cp.on[RAV2D_PASS_DEBLOCK_V] || cp.on[RAV2D_PASS_DEBLOCK_H])) {
            // one tile pass from the first deblocking pass's source to the last one's output
            const bool v = cp.on[RAV2D_PASS_DEBLOCK_V], h = cp.on[RAV2D_PASS_DEBLOCK_H];
            DeblockVHArgs va;
            va.a = v ? cp.dbv : cp.dbh;
            for (int pl = 0; pl < 3; pl++) {
                va.dst[pl] = h ? cp.dbh.dst[pl] : cp.dbv.dst[pl];
                va.on_v[pl] = v && cp.dbv.plane_on[pl];
                va.on_h[pl] = h && cp.dbh.plane_on[pl];
            }
            va.map = deblock_vh_tile_map(va.a, deblock_rows);
            hb.dbvh[g].push_back(va);
            add(RAV2D_PASS_DEBLOCK_V, va.map.first[3]);
        } else if (cp.on[RAV2D_PASS_DEBLOCK_V]) {
            hb.dbv[g].push_back(cp.dbv);
            add(RAV2D_PASS_DEBLOCK_V, cp.dbv.map.first[3]);
        }
        if (!deblock_rows && cp.on[RAV2D_PASS_DEBLOCK_H]) {
            hb.dbh[g].push_back(cp.dbh);
            add(RAV2D_PASS_DEBLOCK_H, cp.dbh.map.first[3]);
        }
        if (cp.on[RAV2D_PASS_CDEF]) {
            hb.cdef[g].push_back(cp.cdef);
            add(RAV2D_PASS_CDEF, cdef_items(cp.cdef));
        }
        hb.lr[g].push_back(cp.lr);
        add(RAV2D_PASS_LR, cp.n_lr);
        if (apply_grain) {
            GrainArgs ga;
            if (frame_has_grain(f->params)) {
                r = grain_args(f, ga);
                if (r) return r;
                hb.gsetup.push_back(ga);
            } else {
                memset(&ga, 0, sizeof(ga));
                ga.bitdepth = f->g.bitdepth;
                for (int pl = 0; pl < 3; pl++) {
                    ga.src[pl] = plane_of(f, BUF_OUT, pl);
                    ga.dst[pl] = plane_of(f, BUF_DISP, pl);
                }
            }
            grain_tile_map(ga);
            hb.gapply[g].push_back(ga);
            add(RAV2D_PASS_GRAIN, ga.map.first[3]);
        }
    }
    for (int p = 0; p < kPasses; p++)
        for (int g = 0; g < kDepths; g++)
            if (hb.total[p][g] > INT32_MAX) {
                set_last_error("batch too large: more than 2^31 work items in one launch");
                return RAV2D_EINVAL;
            }
    return RAV2D_OK;
}

size_t packed_size(const HostBatch &hb) { return layout_of(hb).total; }

size_t packed_bound(int n) {
    // per frame at most: both deblock passes or the fused one, CDEF, LR, grain apply and set-up,
    // one prefix entry per pass; each array (5 + kPasses per bit-depth group,
    // and the set-ups) pads to 256 bytes
    const size_t per = 2 * sizeof(DeblockArgs) + sizeof(DeblockVHArgs) + sizeof(CdefArgs) + sizeof(LrArgs) + 2 * sizeof(GrainArgs) +
                       kPasses * sizeof(int32_t);
    return (size_t)std::max(n, 1) * per + (kDepths * (6 + kPasses) + 1) * 256 + 256;
}

void pack_batch(const HostBatch &hb, uint8_t *host, uint8_t *dev, DevBatch &db) {
    const Layout L = layout_of(hb);
    auto put = [&](size_t o, const void *src, size_t bytes) {
        if (bytes) memcpy(host + o, src, bytes);
    };
    db = DevBatch();
    db.deblock_rows = hb.deblock_rows;
    db.any_ss = hb.any_ss;
    for (int g = 0; g < kDepths; g++) {
        put(L.dbv[g], hb.dbv[g].data(), hb.dbv[g].size() * sizeof(DeblockArgs));
        put(L.dbh[g], hb.dbh[g].data(), hb.dbh[g].size() * sizeof(DeblockArgs));
        put(L.dbvh[g], hb.dbvh[g].data(), hb.dbvh[g].size() * sizeof(DeblockVHArgs));
        put(L.cdef[g], hb.cdef[g].data(), hb.cdef[g].size() * sizeof(CdefArgs));
        put(L.lr[g], hb.lr[g].data(), hb.lr[g].size() * sizeof(LrArgs));
        put(L.gapply[g], hb.gapply[g].data(), hb.gapply[g].size() * sizeof(GrainArgs));
        db.dbv[g] = (DeblockArgs *)(dev + L.dbv[g]);
        db.dbh[g] = (DeblockArgs *)(dev + L.dbh[g]);
        db.dbvh[g] = (DeblockVHArgs *)(dev + L.dbvh[g]);
        db.cdef[g] = (CdefArgs *)(dev + L.cdef[g]);
        db.lr[g] = (LrArgs *)(dev + L.lr[g]);
        db.gapply[g] = (GrainArgs *)(dev + L.gapply[g]);
        for (int p = 0; p < kPasses; p++) {
            put(L.prefix[p][g], hb.prefix[p][g].data(), hb.prefix[p][g].size() * sizeof(int32_t));
            db.prefix[p][g] = (int32_t *)(dev + L.prefix[p][g]);
            db.nf[p][g] = (int)hb.prefix[p][g].size();
            db.total[p][g] = (int32_t)hb.total[p][g];
        }
    }
    put(L.gsetup, hb.gsetup.data(), hb.gsetup.size() * sizeof(GrainArgs));
    db.gsetup = (GrainArgs *)(dev + L.gsetup);
    db.n_setup = (int)hb.gsetup.size();
}

int enqueue_batch(const DevBatch &db, const LaunchShape &shape, bool apply_grain, const EnqueueCtx &ec) {
    if (injected_failure(RAV2D_FAIL_LAUNCH)) {
        set_last_error("batch launch failed (injected)");
        return RAV2D_EDEVICE;
    }
    const hipStream_t s = ec.s;
    const bool grain = apply_grain && db.items(RAV2D_PASS_GRAIN) > 0;
    // the 8-bit group and the high-bit-depth groups are independent chains: with
    // a second stream they run side by side and fill each other's launch gaps and tails
    bool both = false;
    for (int p = 0; p < kPasses; p++) both = both || (db.total[p][0] > 0 && db.items(p) > db.total[p][0]);
    const bool par = both && !shape.mixed_depths && ec.side2 && ec.time_pass < 0 && !ec.profile;
    if (par) {
        HIP_TRY(hipEventRecord(ec.ev_fork2, s));
        HIP_TRY(hipStreamWaitEvent(ec.side2, ec.ev_fork2, 0));
    }
    if (grain && db.n_setup) {
        HIP_TRY(hipEventRecord(ec.ev_fork, s));
        HIP_TRY(hipStreamWaitEvent(ec.side, ec.ev_fork, 0));
        launch_grain_setup_batch(db.gsetup, db.n_setup, ec.side);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(ec.ev_join, ec.side));
    }
    for (int p = 0; p < kPasses; p++) {
        if (db.items(p) == 0) continue;
        if (p == RAV2D_PASS_GRAIN && !grain) continue;
        if (p == RAV2D_PASS_GRAIN && db.n_setup) {
            HIP_TRY(hipStreamWaitEvent(s, ec.ev_join, 0));
            if (par) HIP_TRY(hipStreamWaitEvent(ec.side2, ec.ev_join, 0));
        }
        if (p == ec.time_pass) HIP_TRY(hipEventRecord(ec.ev_a, s));
        if (shape.mixed_depths) {
            // one launch over every group's items
            PassTimer pt(ec.profile, p, s);
            MixSched ms;
            ms.first[0] = 0;
            for (int g = 0; g < kDepths; g++) {
                ms.g[g].prefix = db.prefix[p][g];
                ms.g[g].n_frames = db.nf[p][g];
                ms.g[g].total = db.total[p][g];
                ms.g[g].xcd_remap = 0;
                ms.first[g + 1] = ms.first[g] + db.total[p][g];
            }
            ms.xcd_remap = shape.xcd_remap[p] != 0;
            const int grid = grid_of(ms.first[3], shape, p);
            switch (p) {
                case RAV2D_PASS_DEBLOCK_V:
                    if (db.deblock_rows)
                        launch_deblock_vh_mix({{db.dbvh[0], db.dbvh[1], db.dbvh[2]}}, ms, grid, db.deblock_rows,
                                              db.any_ss, s);
                    else
                        launch_deblock_mix(true, {{db.dbv[0], db.dbv[1], db.dbv[2]}}, ms, grid, db.any_ss, s);
                    break;
                case RAV2D_PASS_DEBLOCK_H:
                    launch_deblock_mix(false, {{db.dbh[0], db.dbh[1], db.dbh[2]}}, ms, grid, db.any_ss, s);
                    break;
                case RAV2D_PASS_CDEF: launch_cdef_mix({{db.cdef[0], db.cdef[1], db.cdef[2]}}, ms, grid, db.any_ss, s); break;
                case RAV2D_PASS_LR: launch_lr_mix({{db.lr[0], db.lr[1], db.lr[2]}}, ms, grid, s); break;
                default: launch_grain_apply_mix({{db.gapply[0], db.gapply[1], db.gapply[2]}}, ms, grid, db.any_ss, s); break;
            }
            HIP_TRY(hipGetLastError());
            if (p == ec.time_pass) HIP_TRY(hipEventRecord(ec.ev_b, s));
            continue;
        }
        for (int g = 0; g < kDepths; g++) {
            if (db.total[p][g] == 0) continue;
            const hipStream_t sg = (par && g == 0) ? ec.side2 : s;
            PassTimer pt(ec.profile, p, sg);
            BatchSched sc;
            sc.prefix = db.prefix[p][g];
            sc.n_frames = db.nf[p][g];
            sc.total = db.total[p][g];
            sc.xcd_remap = shape.xcd_remap[p] != 0;
            const int grid = grid_of(db.total[p][g], shape, p);
            const int bd = group_depth(g);
            switch (p) {
                case RAV2D_PASS_DEBLOCK_V:
                    if (db.deblock_rows)
                        launch_deblock_vh_batch(db.dbvh[g], sc, grid, bd, db.deblock_rows, db.any_ss, sg);
                    else
                        launch_deblock_batch(true, db.dbv[g], sc, grid, bd, db.any_ss, sg);
                    break;
                case RAV2D_PASS_DEBLOCK_H: launch_deblock_batch(false, db.dbh[g], sc, grid, bd, db.any_ss, sg); break;
                case RAV2D_PASS_CDEF: launch_cdef_batch(db.cdef[g], sc, grid, bd, db.any_ss, sg); break;
                case RAV2D_PASS_LR: launch_lr_batch(db.lr[g], sc, grid, bd, sg); break;
                default: launch_grain_apply_batch(db.gapply[g], sc, grid, bd, db.any_ss, sg); break;
            }
            HIP_TRY(hipGetLastError());
        }
        if (p == ec.time_pass) HIP_TRY(hipEventRecord(ec.ev_b, s));
    }
    if (par) {
        HIP_TRY(hipEventRecord(ec.ev_join2, ec.side2));
        HIP_TRY(hipStreamWaitEvent(s, ec.ev_join2, 0));
    }
    return RAV2D_OK;
}

}  // namespace r2g

using namespace r2g;

struct Rav2dGpuBatch {
    Rav2dGpuCtx *ctx = nullptr;
    std::vector<Rav2dGpuFrame *> frames;
    std::vector<uint64_t> gens;
    Rav2dBatchConfig cfg;
    HostBatch hb;
    DevBatch db;
    uint8_t *dmem = nullptr;  // packed argument block
    size_t dbytes = 0;
    hipStream_t side = nullptr, side2 = nullptr, cap = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_fork2 = nullptr, ev_join2 = nullptr, ev_a = nullptr,
               ev_b = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void drop_graph(Rav2dGpuBatch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void drop_args(Rav2dGpuBatch *b) {
    if (b->dmem) {
        (void)hipFree(b->dmem);
        std::lock_guard<std::mutex> lk(b->ctx->mu);
        b->ctx->in_use -= b->dbytes;
    }
    b->dmem = nullptr;
    b->dbytes = 0;
}

LaunchShape shape_of(const Rav2dBatchConfig &c) {
    LaunchShape sh;
    for (int p = 0; p < kPasses; p++) {
        sh.grid[p] = c.grid[p];
        sh.xcd_remap[p] = c.xcd_remap[p];
    }
    sh.mixed_depths = c.mixed_depths;
    return sh;
}

// Plans every frame and uploads the argument block.
int build(Rav2dGpuBatch *b) {
    drop_graph(b);
    drop_args(b);
    int r = batch_init_device();
    if (r) return r;
    r = plan_batch(b->ctx, b->frames.data(), (int)b->frames.size(), b->cfg.apply_grain != 0, b->hb,
                   b->cfg.deblock_rows);
    if (r) return r;
    const size_t bytes = packed_size(b->hb);
    {
        std::lock_guard<std::mutex> lk(b->ctx->mu);
        if (b->ctx->budget && b->ctx->in_use + bytes > b->ctx->budget) {
            set_last_error("HBM budget exceeded");
            return RAV2D_EAGAIN;
        }
        b->ctx->in_use += bytes;
    }
    if (injected_failure(RAV2D_FAIL_ALLOC) || hipMalloc((void **)&b->dmem, bytes) != hipSuccess) {
        std::lock_guard<std::mutex> lk(b->ctx->mu);
        b->ctx->in_use -= bytes;
        b->dmem = nullptr;
        set_last_error("hipMalloc failed");
        return RAV2D_ENOMEM;
    }
    b->dbytes = bytes;
    std::vector<uint8_t> host(bytes, 0);
    pack_batch(b->hb, host.data(), b->dmem, b->db);
    HIP_TRY(hipMemcpy(b->dmem, host.data(), bytes, hipMemcpyHostToDevice));
    for (size_t i = 0; i < b->frames.size(); i++) b->gens[i] = b->frames[i]->gen;
    return RAV2D_OK;
}

int enqueue(Rav2dGpuBatch *b, hipStream_t s, int time_pass, bool timers) {
    EnqueueCtx ec;
    ec.s = s;
    ec.side = b->side;
    ec.ev_fork = b->ev_fork;
    ec.ev_join = b->ev_join;
    ec.side2 = b->side2;
    ec.ev_fork2 = b->ev_fork2;
    ec.ev_join2 = b->ev_join2;
    ec.time_pass = time_pass;
    ec.ev_a = b->ev_a;
    ec.ev_b = b->ev_b;
    ec.profile = timers ? b->ctx : nullptr;
    return enqueue_batch(b->db, shape_of(b->cfg), b->cfg.apply_grain != 0, ec);
}

int capture(Rav2dGpuBatch *b) {
    drop_graph(b);
    HIP_TRY(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    int r = enqueue(b, b->cap, -1, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) {
        set_last_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        return RAV2D_EDEVICE;
    }
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return RAV2D_OK;
}

bool stale(const Rav2dGpuBatch *b) {
    for (size_t i = 0; i < b->frames.size(); i++)
        if (b->frames[i]->gen != b->gens[i]) return true;
    return false;
}

int prepare(Rav2dGpuBatch *b) {
    if (!b->dmem || stale(b)) {
        int r = build(b);
        if (r) return r;
    }
    if (b->cfg.use_graph && !b->exec) return capture(b);
    return RAV2D_OK;
}

bool config_ok(const Rav2dBatchConfig &c) {
    for (int p = 0; p < kPasses; p++)
        if (c.grid[p] < 0) return false;
    if (c.deblock_rows && !deblock_vh_rows_ok(c.deblock_rows)) {
        set_last_error("batch config: deblock_rows must be 0, 32 or 64");
        return false;
    }
    if (c.mixed_depths != 0 && c.mixed_depths != 1) {
        set_last_error("batch config: mixed_depths must be 0 or 1");
        return false;
    }
    return true;
}

}  // namespace

extern "C" {

void rav2d_gpu_batch_default_config(Rav2dBatchConfig *cfg) {
    if (!cfg) return;
    memset(cfg, 0, sizeof(*cfg));
    cfg->use_graph = 1;
    cfg->apply_grain = 1;
}

int rav2d_gpu_batch_create(Rav2dGpuCtx *ctx, Rav2dGpuFrame *const *frames, int n, const Rav2dBatchConfig *cfg,
                           Rav2dGpuBatch **out) {
    if (!ctx || !frames || n <= 0 || !out) return RAV2D_EINVAL;
    *out = nullptr;
    for (int i = 0; i < n; i++)
        if (!frames[i] || frames[i]->ctx != ctx) {
            set_last_error("batch frames must belong to the context");
            return RAV2D_EINVAL;
        }
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (frames[i] == frames[j]) {
                set_last_error("a frame may appear only once in a batch");
                return RAV2D_EINVAL;
            }
    HIP_TRY(hipSetDevice(ctx->device));
    Rav2dGpuBatch *b = new Rav2dGpuBatch();
    b->ctx = ctx;
    b->frames.assign(frames, frames + n);
    b->gens.assign(n, 0);
    if (cfg)
        b->cfg = *cfg;
    else
        rav2d_gpu_batch_default_config(&b->cfg);
    if (!config_ok(b->cfg)) {
        delete b;
        return RAV2D_EINVAL;
    }
    if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->side2, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork, kStreamSyncEventFlags) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join, kStreamSyncEventFlags) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork2, kStreamSyncEventFlags) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join2, kStreamSyncEventFlags) != hipSuccess ||
        hipEventCreate(&b->ev_a) != hipSuccess || hipEventCreate(&b->ev_b) != hipSuccess) {
        set_last_error("stream/event creation failed");
        rav2d_gpu_batch_destroy(b);
        return RAV2D_EDEVICE;
    }
    int r = prepare(b);
    if (r) {
        rav2d_gpu_batch_destroy(b);
        return r;
    }
    *out = b;
    return RAV2D_OK;
}

int rav2d_gpu_batch_destroy(Rav2dGpuBatch *b) {
    if (!b) return RAV2D_EINVAL;
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();  // a replay may be in flight on any stream
    drop_graph(b);
    drop_args(b);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->side2) (void)hipStreamDestroy(b->side2);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    for (hipEvent_t e : {b->ev_fork, b->ev_join, b->ev_fork2, b->ev_join2, b->ev_a, b->ev_b})
        if (e) (void)hipEventDestroy(e);
    delete b;
    return RAV2D_OK;
}

int rav2d_gpu_batch_run(Rav2dGpuBatch *b, void *stream) {
    if (!b) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(b->ctx->device));
    int r = prepare(b);
    if (r) return r;
    hipStream_t s = pick_stream(b->ctx, stream);
    if (b->cfg.use_graph) {
        HIP_TRY(hipGraphLaunch(b->exec, s));
        return RAV2D_OK;
    }
    return enqueue(b, s, -1, true);
}

int rav2d_gpu_batch_get_config(const Rav2dGpuBatch *b, Rav2dBatchConfig *out) {
    if (!b || !out) return RAV2D_EINVAL;
    *out = b->cfg;
    return RAV2D_OK;
}

int rav2d_gpu_batch_set_config(Rav2dGpuBatch *b, const Rav2dBatchConfig *cfg) {
    if (!b || !cfg) return RAV2D_EINVAL;
    if (!config_ok(*cfg)) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    bool rebuild = cfg->apply_grain != b->cfg.apply_grain || cfg->deblock_rows != b->cfg.deblock_rows;
    b->cfg = *cfg;
    drop_graph(b);
    if (rebuild) drop_args(b);
    return prepare(b);
}

int rav2d_gpu_batch_info(const Rav2dGpuBatch *b, int pass, int64_t *items, int64_t *bytes, int *frames) {
    if (!b || pass < 0 || pass >= kPasses) return RAV2D_EINVAL;
    if (items) *items = b->hb.items(pass);
    if (bytes) *bytes = b->hb.bytes[pass];
    if (frames) *frames = b->hb.frames[pass];
    return RAV2D_OK;
}

int64_t rav2d_gpu_batch_pixels(const Rav2dGpuBatch *b) { return b ? b->hb.pixels : 0; }

int rav2d_gpu_batch_autotune(Rav2dGpuBatch *b, int iters, Rav2dBatchConfig *chosen) {
    if (!b || iters <= 0) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(b->ctx->device));
    int r = prepare(b);
    if (r) return r;
    hipStream_t s = b->ctx->stream;
    Rav2dBatchConfig best = b->cfg;
    // Times pass p under every launch shape with the rest of `best` (the
    // batch's arguments must be built for best's deblocking form); keeps the
    // fastest shape in best and returns its time (sum over iters runs).
    auto tune_pass = [&](int p, double &best_ms) -> int {
        best_ms = 0;
        if (b->hb.items(p) == 0) return RAV2D_OK;
        std::vector<std::pair<int, int>> cands = {{0, 0}, {0, 1}};
        for (int g : {256, 512, 1024, 2048, 4096})
            if (g < b->hb.items(p)) {
                cands.push_back({g, 0});
                cands.push_back({g, 1});
            }
        best_ms = 1e30;
        for (auto c : cands) {
            b->cfg = best;
            b->cfg.grid[p] = c.first;
            b->cfg.xcd_remap[p] = c.second;
            int e = enqueue(b, s, -1, false);  // warm-up
            if (e) return e;
            double sum = 0;
            for (int it = 0; it < iters; it++) {
                e = enqueue(b, s, p, false);
                if (e) return e;
                HIP_TRY(hipEventSynchronize(b->ev_b));
                float ms = 0.f;
                HIP_TRY(hipEventElapsedTime(&ms, b->ev_a, b->ev_b));
                sum += ms;
            }
            if (sum < best_ms) {
                best_ms = sum;
                best.grid[p] = c.first;
                best.xcd_remap[p] = c.second;
            }
        }
        return RAV2D_OK;
    };
    // re-plans the batch for configuration c (its deblocking form), eager launches
    auto replan = [&](const Rav2dBatchConfig &c) -> int {
        HIP_TRY(hipStreamSynchronize(s));
        b->cfg = c;
        return build(b);
    };
    // deblocking: the two separate passes, or one tile pass of 32 or 64 rows
    Rav2dBatchConfig sep = best;
    sep.deblock_rows = 0;
    r = replan(sep);
    if (r) return r;
    best = sep;
    double tv = 0, th = 0;
    if ((r = tune_pass(RAV2D_PASS_DEBLOCK_V, tv)) || (r = tune_pass(RAV2D_PASS_DEBLOCK_H, th))) return r;
    double best_db = tv + th;
    Rav2dBatchConfig best_cfg = best;
    for (int rows : {32, 64}) {
        Rav2dBatchConfig fc = best_cfg;
        fc.deblock_rows = rows;
        fc.grid[RAV2D_PASS_DEBLOCK_V] = fc.xcd_remap[RAV2D_PASS_DEBLOCK_V] = 0;
        fc.grid[RAV2D_PASS_DEBLOCK_H] = fc.xcd_remap[RAV2D_PASS_DEBLOCK_H] = 0;
        r = replan(fc);
        if (r) return r;
        best = fc;
        double tf = 0;
        if ((r = tune_pass(RAV2D_PASS_DEBLOCK_V, tf))) return r;
        if (tf < best_db) {
            best_db = tf;
            best_cfg = best;
        }
    }
    r = replan(best_cfg);
    if (r) return r;
    best = best_cfg;
    best.mixed_depths = 0;
    for (int p = RAV2D_PASS_CDEF; p < kPasses; p++) {
        if (p == RAV2D_PASS_GRAIN && !b->cfg.apply_grain) continue;
        double t = 0;
        if ((r = tune_pass(p, t))) return r;
    }
    // one launch per pass over every bit depth (shapes tuned again for it)
    // against one launch per bit depth on parallel branches: whole-batch time
    bool several = false;
    for (int p = 0; p < kPasses; p++) {
        int groups = 0;
        for (int g = 0; g < kDepths; g++) groups += b->hb.total[p][g] > 0;
        several = several || groups > 1;
    }
    if (several) {
        auto whole = [&](const Rav2dBatchConfig &c, double &ms_out) -> int {
            b->cfg = c;
            int e = enqueue(b, s, -1, false);  // warm-up
            if (e) return e;
            ms_out = 0;
            for (int it = 0; it < iters; it++) {
                HIP_TRY(hipEventRecord(b->ev_a, s));
                e = enqueue(b, s, -1, false);
                if (e) return e;
                HIP_TRY(hipEventRecord(b->ev_b, s));
                HIP_TRY(hipEventSynchronize(b->ev_b));
                float ms = 0.f;
                HIP_TRY(hipEventElapsedTime(&ms, b->ev_a, b->ev_b));
                ms_out += ms;
            }
            return RAV2D_OK;
        };
        const Rav2dBatchConfig split = best;
        best.mixed_depths = 1;
        for (int p = 0; p < kPasses; p++) {
            if (p == RAV2D_PASS_GRAIN && !b->cfg.apply_grain) continue;
            double t = 0;
            if ((r = tune_pass(p, t))) return r;
        }
        const Rav2dBatchConfig mixed = best;
        double t_split = 0, t_mixed = 0;
        if ((r = whole(split, t_split)) || (r = whole(mixed, t_mixed))) return r;
        best = t_mixed < t_split ? mixed : split;
    }
    HIP_TRY(hipStreamSynchronize(s));
    b->cfg = best;
    drop_graph(b);
    r = prepare(b);
    if (r) return r;
    if (chosen) *chosen = best;
    return RAV2D_OK;
}

}  // extern "C"
