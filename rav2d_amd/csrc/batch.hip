// Batched filtering: many frames of any size and bit depth, one launch per
// pass over all of them, optionally captured once and replayed as a hipGraph.
//
// Each pass's launch gets an array of per-frame argument blocks and the prefix
// sum of the frames' work items (tiles, superblocks, restoration units); a
// workgroup finds its frame by binary search over the prefix and dispatches on
// the frame's bit depth at run time.  The launch shape (one workgroup per item,
// or a persistent grid that loops; XCD-contiguous item ranges or round robin)
// is per pass and is what rav2d_gpu_batch_autotune picks.
//
// Pass order per frame is that of the per-frame chain (plan_chain): deblock V,
// deblock H, CDEF/CCSO, loop restoration into the output planes, film grain
// into the display planes.  LR's launch also carries copy items for every
// plane it does not restore, so one launch leaves every frame's output planes
// complete; grain's launch copies the planes of frames without grain.  The
// grain tables of all grained frames are built by one set-up launch on a side
// stream that forks at the start of the batch and joins before grain.
#include <string.h>

#include <algorithm>
#include <vector>

#include "frame.h"

using namespace r2g;

namespace {

constexpr int kPasses = RAV2D_NUM_PASSES;

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

struct Rav2dGpuBatch {
    Rav2dGpuCtx *ctx = nullptr;
    std::vector<Rav2dGpuFrame *> frames;
    std::vector<uint64_t> gens;
    Rav2dBatchConfig cfg;
    // device argument blocks, one allocation
    uint8_t *dmem = nullptr;
    size_t dbytes = 0;
    DeblockArgs *d_dbv = nullptr, *d_dbh = nullptr;
    CdefArgs *d_cdef = nullptr;
    LrArgs *d_lr = nullptr;
    GrainArgs *d_gapply = nullptr, *d_gsetup = nullptr;
    LrItem *d_items = nullptr;
    int32_t *d_prefix[kPasses] = {};
    int nf[kPasses] = {};         // frames in each pass's launch
    int64_t total[kPasses] = {};  // work items of each pass's launch
    int64_t bytes[kPasses] = {};  // algorithmic bytes of each pass (read + write of the planes it covers)
    int n_setup = 0;
    int64_t pixels = 0;           // luma pixels of the batch
    hipStream_t side = nullptr, cap = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_a = nullptr, ev_b = nullptr;
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

// Copy items for plane `pl`: whole-plane copy from the plane's current buffer
// to OUT in 256x64 blocks (LR items of kind NONE without GDF).
void copy_items(const Rav2dGpuFrame *f, int pl, std::vector<LrItem> &out) {
    for (int y = 0; y < f->ph[pl]; y += 64)
        for (int x = 0; x < f->pw[pl]; x += 256) {
            LrItem it;
            memset(&it, 0, sizeof(it));
            it.plane = (int16_t)pl;
            it.x0 = (int16_t)x;
            it.w = (int16_t)std::min(256, f->pw[pl] - x);
            it.y0 = (int16_t)y;
            it.h = (int16_t)std::min(64, f->ph[pl] - y);
            it.kind = LR_KIND_NONE;
            out.push_back(it);
        }
}

int grid_of(const Rav2dGpuBatch *b, int pass) {
    int64_t total = b->total[pass];
    int g = b->cfg.grid[pass];
    int64_t n = g > 0 ? std::min<int64_t>(g, total) : total;
    if (b->cfg.xcd_remap[pass]) n = (n + 7) & ~int64_t(7);
    return (int)std::max<int64_t>(n, 1);
}

BatchSched sched_of(const Rav2dGpuBatch *b, int pass) {
    BatchSched s;
    s.prefix = b->d_prefix[pass];
    s.n_frames = b->nf[pass];
    s.total = (int32_t)b->total[pass];
    s.xcd_remap = b->cfg.xcd_remap[pass] != 0;
    return s;
}

// Plans every frame and uploads the argument blocks.
int build(Rav2dGpuBatch *b) {
    drop_graph(b);
    drop_args(b);
    const int n = (int)b->frames.size();
    std::vector<DeblockArgs> dbv, dbh;
    std::vector<CdefArgs> cdef;
    std::vector<LrArgs> lr;
    std::vector<GrainArgs> gapply, gsetup;
    std::vector<int32_t> prefix[kPasses];
    std::vector<LrItem> items;
    std::vector<size_t> lr_off;
    for (int p = 0; p < kPasses; p++) {
        b->total[p] = 0;
        b->bytes[p] = 0;
    }
    b->pixels = 0;
    cdef_init_consts();
    if (!grain_init_tables()) {
        set_last_error("film grain tables: device allocation failed");
        return RAV2D_ENOMEM;
    }
    for (int i = 0; i < n; i++) {
        Rav2dGpuFrame *f = b->frames[i];
        ChainPlan cp;
        int r = plan_chain(b->ctx, f, RAV2D_PASS_LR, cp);
        if (r) return r;
        // one read and one write of the frame's 1.5 * W * H samples
        const int64_t fbytes = 3 * (int64_t)f->pw[0] * f->ph[0] * f->bps;
        b->pixels += (int64_t)f->pw[0] * f->ph[0];
        auto add = [&](int pass, int64_t count) {
            prefix[pass].push_back((int32_t)b->total[pass]);
            b->total[pass] += count;
            b->bytes[pass] += fbytes;
        };
        if (cp.on[RAV2D_PASS_DEBLOCK_V]) {
            dbv.push_back(cp.dbv);
            add(RAV2D_PASS_DEBLOCK_V, cp.dbv.map.first[3]);
        }
        if (cp.on[RAV2D_PASS_DEBLOCK_H]) {
            dbh.push_back(cp.dbh);
            add(RAV2D_PASS_DEBLOCK_H, cp.dbh.map.first[3]);
        }
        if (cp.on[RAV2D_PASS_CDEF]) {
            cdef.push_back(cp.cdef);
            add(RAV2D_PASS_CDEF, cdef_items(cp.cdef));
        }
        // LR: restoration items, then plane copies for what LR leaves alone
        lr_off.push_back(items.size());
        size_t first = items.size();
        if (cp.on[RAV2D_PASS_LR]) items.insert(items.end(), f->lr_items.begin(), f->lr_items.end());
        for (int pl = 0; pl < 3; pl++)
            if (cp.final_buf[pl] != BUF_OUT) copy_items(f, pl, items);
        lr.push_back(cp.lr);
        add(RAV2D_PASS_LR, (int64_t)(items.size() - first));
        if (b->cfg.apply_grain) {
            GrainArgs g;
            if (frame_has_grain(f->params)) {
                r = grain_args(f, g);
                if (r) return r;
                gsetup.push_back(g);
            } else {
                memset(&g, 0, sizeof(g));
                g.bitdepth = f->g.bitdepth;
                for (int pl = 0; pl < 3; pl++) {
                    g.src[pl] = plane_of(f, BUF_OUT, pl);
                    g.dst[pl] = plane_of(f, BUF_DISP, pl);
                }
            }
            grain_tile_map(g);
            gapply.push_back(g);
            add(RAV2D_PASS_GRAIN, g.map.first[3]);
        }
    }
    for (int p = 0; p < kPasses; p++) {
        b->nf[p] = (int)prefix[p].size();
        if (b->total[p] > INT32_MAX) {
            set_last_error("batch too large: more than 2^31 work items in one pass");
            return RAV2D_EINVAL;
        }
    }
    b->n_setup = (int)gsetup.size();
    // one device block: argument arrays, prefixes, LR items
    size_t off = 0;
    auto reserve = [&](size_t bytes) {
        size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    size_t o_dbv = reserve(dbv.size() * sizeof(DeblockArgs));
    size_t o_dbh = reserve(dbh.size() * sizeof(DeblockArgs));
    size_t o_cdef = reserve(cdef.size() * sizeof(CdefArgs));
    size_t o_lr = reserve(lr.size() * sizeof(LrArgs));
    size_t o_gapply = reserve(gapply.size() * sizeof(GrainArgs));
    size_t o_gsetup = reserve(gsetup.size() * sizeof(GrainArgs));
    size_t o_items = reserve(items.size() * sizeof(LrItem));
    size_t o_prefix[kPasses];
    for (int p = 0; p < kPasses; p++) o_prefix[p] = reserve(prefix[p].size() * sizeof(int32_t));
    if (off == 0) off = 256;
    {
        std::lock_guard<std::mutex> lk(b->ctx->mu);
        if (b->ctx->budget && b->ctx->in_use + off > b->ctx->budget) {
            set_last_error("HBM budget exceeded");
            return RAV2D_EAGAIN;
        }
        b->ctx->in_use += off;
    }
    if (hipMalloc((void **)&b->dmem, off) != hipSuccess) {
        std::lock_guard<std::mutex> lk(b->ctx->mu);
        b->ctx->in_use -= off;
        b->dmem = nullptr;
        set_last_error("hipMalloc failed");
        return RAV2D_ENOMEM;
    }
    b->dbytes = off;
    b->d_dbv = (DeblockArgs *)(b->dmem + o_dbv);
    b->d_dbh = (DeblockArgs *)(b->dmem + o_dbh);
    b->d_cdef = (CdefArgs *)(b->dmem + o_cdef);
    b->d_lr = (LrArgs *)(b->dmem + o_lr);
    b->d_gapply = (GrainArgs *)(b->dmem + o_gapply);
    b->d_gsetup = (GrainArgs *)(b->dmem + o_gsetup);
    b->d_items = (LrItem *)(b->dmem + o_items);
    for (int p = 0; p < kPasses; p++) b->d_prefix[p] = (int32_t *)(b->dmem + o_prefix[p]);
    for (size_t i = 0; i < lr.size(); i++) lr[i].items = b->d_items + lr_off[i];
    std::vector<uint8_t> host(off, 0);
    auto put = [&](size_t o, const void *src, size_t bytes) {
        if (bytes) memcpy(host.data() + o, src, bytes);
    };
    put(o_dbv, dbv.data(), dbv.size() * sizeof(DeblockArgs));
    put(o_dbh, dbh.data(), dbh.size() * sizeof(DeblockArgs));
    put(o_cdef, cdef.data(), cdef.size() * sizeof(CdefArgs));
    put(o_lr, lr.data(), lr.size() * sizeof(LrArgs));
    put(o_gapply, gapply.data(), gapply.size() * sizeof(GrainArgs));
    put(o_gsetup, gsetup.data(), gsetup.size() * sizeof(GrainArgs));
    put(o_items, items.data(), items.size() * sizeof(LrItem));
    for (int p = 0; p < kPasses; p++) put(o_prefix[p], prefix[p].data(), prefix[p].size() * sizeof(int32_t));
    HIP_TRY(hipMemcpy(b->dmem, host.data(), off, hipMemcpyHostToDevice));
    for (int i = 0; i < n; i++) b->gens[i] = b->frames[i]->gen;
    return RAV2D_OK;
}

// Enqueues the whole batch on `s`.  With `time_pass` >= 0, ev_a / ev_b bracket
// that pass's launch.  `timers` brackets every pass with the context's
// profiling events (eager runs on the context stream only).
int enqueue(Rav2dGpuBatch *b, hipStream_t s, int time_pass, bool timers) {
    const bool grain = b->cfg.apply_grain && b->total[RAV2D_PASS_GRAIN] > 0;
    if (grain && b->n_setup) {
        HIP_TRY(hipEventRecord(b->ev_fork, s));
        HIP_TRY(hipStreamWaitEvent(b->side, b->ev_fork, 0));
        launch_grain_setup_batch(b->d_gsetup, b->n_setup, b->side);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(b->ev_join, b->side));
    }
    for (int p = 0; p < kPasses; p++) {
        if (b->total[p] == 0) continue;
        if (p == RAV2D_PASS_GRAIN && !grain) continue;
        if (p == RAV2D_PASS_GRAIN && b->n_setup) HIP_TRY(hipStreamWaitEvent(s, b->ev_join, 0));
        if (p == time_pass) HIP_TRY(hipEventRecord(b->ev_a, s));
        {
            Rav2dGpuCtx *tctx = timers ? b->ctx : nullptr;
            bool on = tctx && tctx->profiling && s == tctx->stream;
            hipEvent_t e0 = nullptr;
            if (on) {
                e0 = take_event(tctx);
                if (e0) (void)hipEventRecord(e0, s);
            }
            const BatchSched sc = sched_of(b, p);
            const int grid = grid_of(b, p);
            switch (p) {
                case RAV2D_PASS_DEBLOCK_V: launch_deblock_batch(true, b->d_dbv, sc, grid, s); break;
                case RAV2D_PASS_DEBLOCK_H: launch_deblock_batch(false, b->d_dbh, sc, grid, s); break;
                case RAV2D_PASS_CDEF: launch_cdef_batch(b->d_cdef, sc, grid, s); break;
                case RAV2D_PASS_LR: launch_lr_batch(b->d_lr, sc, grid, s); break;
                default: launch_grain_apply_batch(b->d_gapply, sc, grid, s); break;
            }
            HIP_TRY(hipGetLastError());
            if (e0) {
                hipEvent_t e1 = take_event(tctx);
                if (e1) {
                    (void)hipEventRecord(e1, s);
                    tctx->recs.push_back({p, e0, e1});
                }
            }
        }
        if (p == time_pass) HIP_TRY(hipEventRecord(b->ev_b, s));
    }
    return RAV2D_OK;
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
    if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess ||
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
    if (b->cap) (void)hipStreamDestroy(b->cap);
    for (hipEvent_t e : {b->ev_fork, b->ev_join, b->ev_a, b->ev_b})
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
    for (int p = 0; p < kPasses; p++)
        if (cfg->grid[p] < 0) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    bool rebuild = cfg->apply_grain != b->cfg.apply_grain;
    b->cfg = *cfg;
    drop_graph(b);
    if (rebuild) drop_args(b);
    return prepare(b);
}

int rav2d_gpu_batch_info(const Rav2dGpuBatch *b, int pass, int64_t *items, int64_t *bytes, int *frames) {
    if (!b || pass < 0 || pass >= kPasses) return RAV2D_EINVAL;
    if (items) *items = b->total[pass];
    if (bytes) *bytes = b->bytes[pass];
    if (frames) *frames = b->nf[pass];
    return RAV2D_OK;
}

int64_t rav2d_gpu_batch_pixels(const Rav2dGpuBatch *b) { return b ? b->pixels : 0; }

int rav2d_gpu_batch_autotune(Rav2dGpuBatch *b, int iters, Rav2dBatchConfig *chosen) {
    if (!b || iters <= 0) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(b->ctx->device));
    int r = prepare(b);
    if (r) return r;
    hipStream_t s = b->ctx->stream;
    Rav2dBatchConfig best = b->cfg;
    for (int p = 0; p < kPasses; p++) {
        if (b->total[p] == 0) continue;
        if (p == RAV2D_PASS_GRAIN && !b->cfg.apply_grain) continue;
        std::vector<std::pair<int, int>> cands = {{0, 0}, {0, 1}};
        for (int g : {256, 512, 1024, 2048, 4096})
            if (g < b->total[p]) {
                cands.push_back({g, 0});
                cands.push_back({g, 1});
            }
        double best_ms = 1e30;
        for (auto c : cands) {
            b->cfg = best;
            b->cfg.grid[p] = c.first;
            b->cfg.xcd_remap[p] = c.second;
            r = enqueue(b, s, -1, false);  // warm-up
            if (r) return r;
            double sum = 0;
            for (int it = 0; it < iters; it++) {
                r = enqueue(b, s, p, false);
                if (r) return r;
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
