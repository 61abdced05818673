// Serving: a request loop that takes filter jobs from many streams as they
// arrive and runs them with continuous batching.
//
// A job names a device-resident frame (input planes, side info and parameters
// already set).  A dispatcher thread owns `max_inflight` launch slots, one
// HIP stream each.  Whenever a slot is free and jobs wait, it takes up to
// `max_batch` of them (oldest first, skipping jobs whose frame is still in an
// earlier batch), plans them into one batched launch per pass (batch.h) and
// enqueues that on the slot's stream: batches launched while others run
// overlap on the device.  Completion is detected per slot with an event.
// With `checksum`, a hashing launch after the passes reduces each job's
// output (and display) planes to 64-bit sums, so that a job's result can be
// compared with the same frame filtered alone.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <queue>
#include <random>
#include <thread>
#include <vector>

#include "batch.h"

namespace r2g {

// ---- plane checksums --------------------------------------------------------
// sum over samples of splitmix64(key), key = plane << 60 | y << 40 | x << 20 | v:
// order-independent, so the atomics may add in any order.
struct SumArgs {
    DevPlane pl[2][3];  // [which: output, display][plane]
    int32_t nwhich;
    int32_t hbd;
    unsigned long long *out;  // [2] per frame
    int32_t bands[3];          // 16-row bands per plane
    int32_t first[4];          // prefix of bands over the planes
};

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void checksum_batch(const SumArgs *args, BatchSched s) {
    __shared__ unsigned long long part[4];
    for_each_item<false>(s, [&](int f, int t) {
        const SumArgs &a = args[f];
        const int per = a.first[3];
        const int which = t / per;
        const int r = t - which * per;
        const int pl = r >= a.first[2] ? 2 : (r >= a.first[1] ? 1 : 0);
        const int band = r - a.first[pl];
        const DevPlane &p = a.pl[which][pl];
        unsigned long long acc = 0;
        const int y0 = band * 16, y1 = min(p.h, y0 + 16);
        for (int y = y0; y < y1; y++) {
            const uint8_t *row = p.ptr + (int64_t)y * p.stride;
            for (int x = threadIdx.x; x < p.w; x += 256) {
                unsigned v = a.hbd ? ((const uint16_t *)row)[x] : row[x];
                unsigned long long key = ((unsigned long long)pl << 60) | ((unsigned long long)y << 40) |
                                         ((unsigned long long)x << 20) | v;
                acc += mix64(key);
            }
        }
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) atomicAdd(&a.out[which], part[0] + part[1] + part[2] + part[3]);
        __syncthreads();
    });
}

__global__ void fill_u64(unsigned long long *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

SumArgs sum_args(const Rav2dGpuFrame *f, int which0, int nwhich, unsigned long long *out) {
    SumArgs a;
    memset(&a, 0, sizeof(a));
    for (int w = 0; w < nwhich; w++)
        for (int pl = 0; pl < 3; pl++) a.pl[w][pl] = plane_of(f, which0 + w, pl);
    a.nwhich = nwhich;
    a.hbd = f->bps == 2;
    a.out = out;
    a.first[0] = 0;
    for (int pl = 0; pl < 3; pl++) {
        a.bands[pl] = (f->ph[pl] + 15) / 16;
        a.first[pl + 1] = a.first[pl] + a.bands[pl];
    }
    return a;
}

}  // namespace r2g

using namespace r2g;

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    Rav2dGpuFrame *f;
    uint64_t tag;
    double t_submit, t_launch;
};

struct Slot {
    hipStream_t s = nullptr;
    hipStream_t side = nullptr;  // grain set-ups: the context's auxiliary stream, shared by the slots
    hipEvent_t ev_done = nullptr, ev_fork = nullptr, ev_join = nullptr;
    bool busy = false;
    std::vector<Job> jobs;
    int status = RAV2D_OK;
    uint8_t *d_args = nullptr, *h_args = nullptr;  // packed batch arguments (device, pinned host)
    size_t cap = 0;
    SumArgs *d_sum = nullptr, *h_sum = nullptr;
    unsigned long long *d_out = nullptr, *h_out = nullptr;
    int32_t *d_prefix = nullptr, *h_prefix = nullptr;
    int sum_cap = 0;
};

}  // namespace

struct Rav2dGpuServer {
    Rav2dGpuCtx *ctx = nullptr;
    Rav2dServeConfig cfg;
    Clock::time_point t0;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job> pending;
    std::deque<Rav2dServeResult> done;
    std::vector<Slot> slots;
    int64_t submitted = 0, completed = 0;
    bool stop = false;
    bool paused = false;  // no new batches launch while set (queued jobs wait)

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

bool in_flight(const Rav2dGpuServer *sv, const Rav2dGpuFrame *f) {
    for (const Slot &s : sv->slots)
        if (s.busy)
            for (const Job &j : s.jobs)
                if (j.f == f) return true;
    return false;
}

int grow(Slot &sl, size_t bytes, int frames) {
    if (bytes > sl.cap) {
        if (sl.d_args) (void)hipFree(sl.d_args);
        if (sl.h_args) (void)hipHostFree(sl.h_args);
        sl.d_args = sl.h_args = nullptr;
        sl.cap = 0;
        size_t cap = std::max<size_t>(bytes, 64 << 10);
        HIP_TRY(hipMalloc((void **)&sl.d_args, cap));
        HIP_TRY(hipHostMalloc((void **)&sl.h_args, cap, 0));
        sl.cap = cap;
    }
    if (frames > sl.sum_cap) {
        if (sl.d_sum) (void)hipFree(sl.d_sum);
        if (sl.h_sum) (void)hipHostFree(sl.h_sum);
        if (sl.d_out) (void)hipFree(sl.d_out);
        if (sl.h_out) (void)hipHostFree(sl.h_out);
        if (sl.d_prefix) (void)hipFree(sl.d_prefix);
        if (sl.h_prefix) (void)hipHostFree(sl.h_prefix);
        sl.d_sum = sl.h_sum = nullptr;
        sl.d_out = sl.h_out = nullptr;
        sl.d_prefix = sl.h_prefix = nullptr;
        sl.sum_cap = 0;
        int cap = std::max(frames, 16);
        HIP_TRY(hipMalloc((void **)&sl.d_sum, cap * sizeof(SumArgs)));
        HIP_TRY(hipHostMalloc((void **)&sl.h_sum, cap * sizeof(SumArgs), 0));
        HIP_TRY(hipMalloc((void **)&sl.d_out, 2 * cap * sizeof(unsigned long long)));
        HIP_TRY(hipHostMalloc((void **)&sl.h_out, 2 * cap * sizeof(unsigned long long), 0));
        HIP_TRY(hipMalloc((void **)&sl.d_prefix, cap * sizeof(int32_t)));
        HIP_TRY(hipHostMalloc((void **)&sl.h_prefix, cap * sizeof(int32_t), 0));
        sl.sum_cap = cap;
    }
    return RAV2D_OK;
}

// Plans and enqueues the jobs of `sl` on its stream; records ev_done.
int launch(Rav2dGpuServer *sv, Slot &sl) {
    const int n = (int)sl.jobs.size();
    std::vector<Rav2dGpuFrame *> frames(n);
    for (int i = 0; i < n; i++) frames[i] = sl.jobs[i].f;
    HostBatch hb;
    int r = plan_batch(sv->ctx, frames.data(), n, sv->cfg.apply_grain != 0, hb);
    if (r) return r;
    const size_t bytes = packed_size(hb);
    r = grow(sl, bytes, n);
    if (r) return r;
    DevBatch db;
    pack_batch(hb, sl.h_args, sl.d_args, db);
    HIP_TRY(hipMemcpyAsync(sl.d_args, sl.h_args, bytes, hipMemcpyHostToDevice, sl.s));
    EnqueueCtx ec;
    ec.s = sl.s;
    ec.side = sl.side;
    ec.ev_fork = sl.ev_fork;
    ec.ev_join = sl.ev_join;
    LaunchShape shape;
    for (int p = 0; p < kPasses; p++) {
        shape.grid[p] = 0;
        shape.xcd_remap[p] = 0;
    }
    r = enqueue_batch(db, shape, sv->cfg.apply_grain != 0, ec);
    if (r) return r;
    if (sv->cfg.checksum) {
        const int nwhich = sv->cfg.apply_grain ? 2 : 1;
        int64_t total = 0;
        for (int i = 0; i < n; i++) {
            sl.h_sum[i] = sum_args(frames[i], BUF_OUT, nwhich, sl.d_out + 2 * i);  // OUT, then DISP
            sl.h_prefix[i] = (int32_t)total;
            total += (int64_t)nwhich * sl.h_sum[i].first[3];
        }
        HIP_TRY(hipMemcpyAsync(sl.d_sum, sl.h_sum, n * sizeof(SumArgs), hipMemcpyHostToDevice, sl.s));
        HIP_TRY(hipMemcpyAsync(sl.d_prefix, sl.h_prefix, n * sizeof(int32_t), hipMemcpyHostToDevice, sl.s));
        hipLaunchKernelGGL(fill_u64, dim3((2 * n + 255) / 256), dim3(256), 0, sl.s, sl.d_out, 2 * n);
        BatchSched bs;
        bs.prefix = sl.d_prefix;
        bs.n_frames = n;
        bs.total = (int32_t)total;
        bs.xcd_remap = 0;
        hipLaunchKernelGGL(checksum_batch, dim3((unsigned)total), dim3(256), 0, sl.s, sl.d_sum, bs);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(sl.h_out, sl.d_out, 2 * n * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               sl.s));
    }
    HIP_TRY(hipEventRecord(sl.ev_done, sl.s));
    return RAV2D_OK;
}

// Moves the jobs of finished slots to the done queue.  Caller holds sv->mu.
bool retire(Rav2dGpuServer *sv) {
    bool any = false;
    for (Slot &sl : sv->slots) {
        if (!sl.busy) continue;
        if (sl.status == RAV2D_OK && hipEventQuery(sl.ev_done) == hipErrorNotReady) continue;
        const double t = sv->now();
        for (size_t i = 0; i < sl.jobs.size(); i++) {
            const Job &j = sl.jobs[i];
            Rav2dServeResult res;
            memset(&res, 0, sizeof(res));
            res.tag = j.tag;
            res.t_submit = j.t_submit;
            res.t_launch = j.t_launch;
            res.t_done = t;
            res.status = sl.status;
            res.batch_frames = (int32_t)sl.jobs.size();
            if (sv->cfg.checksum && sl.status == RAV2D_OK) {
                res.checksum[0] = sl.h_out[2 * i];
                res.checksum[1] = sv->cfg.apply_grain ? sl.h_out[2 * i + 1] : 0;
            }
            sv->done.push_back(res);
            sv->completed++;
        }
        sl.jobs.clear();
        sl.busy = false;
        any = true;
    }
    return any;
}

void dispatcher(Rav2dGpuServer *sv) {
    (void)hipSetDevice(sv->ctx->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        if (retire(sv)) sv->cv_done.notify_all();
        bool any_busy = false;
        Slot *free_slot = nullptr;
        for (Slot &sl : sv->slots) {
            if (sl.busy)
                any_busy = true;
            else if (!free_slot)
                free_slot = &sl;
        }
        if (sv->stop && sv->pending.empty() && !any_busy) break;
        // jobs whose frame is idle, oldest first; a frame's later jobs wait
        std::vector<Job> take;
        if (free_slot && !sv->pending.empty() && !(sv->paused && !sv->stop)) {
            std::vector<const Rav2dGpuFrame *> seen;
            for (auto it = sv->pending.begin(); it != sv->pending.end() && (int)take.size() < sv->cfg.max_batch;) {
                const Rav2dGpuFrame *f = it->f;
                bool blocked = in_flight(sv, f) || std::find(seen.begin(), seen.end(), f) != seen.end();
                seen.push_back(f);
                if (blocked) {
                    ++it;
                    continue;
                }
                take.push_back(*it);
                it = sv->pending.erase(it);
            }
        }
        if (!take.empty()) {
            const double t = sv->now();
            for (Job &j : take) j.t_launch = t;
            free_slot->jobs = take;
            free_slot->busy = true;
            lk.unlock();
            int r = launch(sv, *free_slot);
            if (r) (void)hipStreamSynchronize(free_slot->s);  // let a partial enqueue finish
            lk.lock();
            free_slot->status = r;
            continue;
        }
        if (any_busy)
            sv->cv_work.wait_for(lk, std::chrono::microseconds(20));
        else
            sv->cv_work.wait(lk);
    }
}

}  // namespace

extern "C" {

void rav2d_gpu_server_default_config(Rav2dServeConfig *cfg) {
    if (!cfg) return;
    memset(cfg, 0, sizeof(*cfg));
    cfg->max_batch = 16;
    cfg->max_inflight = 2;
    cfg->apply_grain = 1;
    cfg->checksum = 0;
}

int rav2d_gpu_server_create(Rav2dGpuCtx *ctx, const Rav2dServeConfig *cfg, Rav2dGpuServer **out) {
    if (!ctx || !out) return RAV2D_EINVAL;
    *out = nullptr;
    Rav2dServeConfig c;
    if (cfg)
        c = *cfg;
    else
        rav2d_gpu_server_default_config(&c);
    if (c.max_batch < 1 || c.max_batch > 4096 || c.max_inflight < 1 || c.max_inflight > 8) {
        set_last_error("server: max_batch must be in [1, 4096] and max_inflight in [1, 8]");
        return RAV2D_EINVAL;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    int r = batch_init_device();
    if (r) return r;
    Rav2dGpuServer *sv = new Rav2dGpuServer();
    sv->ctx = ctx;
    sv->cfg = c;
    sv->t0 = Clock::now();
    sv->slots.resize(c.max_inflight);
    bool ok = true;
    // one stream per slot plus the context's two: the process stays within the
    // device's hardware queues (4 by default) at the default max_inflight, so
    // no slot's launches queue behind another stream's on a shared queue
    for (Slot &sl : sv->slots) {
        ok = ok && hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking) == hipSuccess;
        sl.side = ctx->aux;
        ok = ok && hipEventCreateWithFlags(&sl.ev_done, hipEventDisableTiming) == hipSuccess;
        ok = ok && hipEventCreateWithFlags(&sl.ev_fork, kStreamSyncEventFlags) == hipSuccess;
        ok = ok && hipEventCreateWithFlags(&sl.ev_join, kStreamSyncEventFlags) == hipSuccess;
    }
    if (!ok) {
        set_last_error("server: stream/event creation failed");
        rav2d_gpu_server_destroy(sv);
        return RAV2D_EDEVICE;
    }
    // argument and checksum buffers for the largest batch, allocated here so
    // the dispatcher never allocates while jobs run
    for (Slot &sl : sv->slots) {
        r = grow(sl, packed_bound(c.max_batch), c.max_batch);
        if (r) {
            rav2d_gpu_server_destroy(sv);
            return r;
        }
    }
    sv->th = std::thread(dispatcher, sv);
    *out = sv;
    return RAV2D_OK;
}

int rav2d_gpu_server_destroy(Rav2dGpuServer *sv) {
    if (!sv) return RAV2D_EINVAL;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->th.joinable()) sv->th.join();
    (void)hipSetDevice(sv->ctx->device);
    (void)hipDeviceSynchronize();
    for (Slot &sl : sv->slots) {
        if (sl.s) (void)hipStreamDestroy(sl.s);
        for (hipEvent_t e : {sl.ev_done, sl.ev_fork, sl.ev_join})
            if (e) (void)hipEventDestroy(e);
        for (void *p : {(void *)sl.d_args, (void *)sl.d_sum, (void *)sl.d_out, (void *)sl.d_prefix})
            if (p) (void)hipFree(p);
        for (void *p : {(void *)sl.h_args, (void *)sl.h_sum, (void *)sl.h_out, (void *)sl.h_prefix})
            if (p) (void)hipHostFree(p);
    }
    delete sv;
    return RAV2D_OK;
}

double rav2d_gpu_server_clock(const Rav2dGpuServer *sv) { return sv ? sv->now() : 0.0; }

int rav2d_gpu_server_submit(Rav2dGpuServer *sv, Rav2dGpuFrame *f, uint64_t tag, double t_arrival) {
    if (!sv || !f || f->ctx != sv->ctx) return RAV2D_EINVAL;
    if (!f->side_ok) {
        set_last_error("side info not set");
        return RAV2D_EINVAL;
    }
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->stop) return RAV2D_EINVAL;
        Job j;
        j.f = f;
        j.tag = tag;
        j.t_submit = t_arrival >= 0 ? t_arrival : sv->now();
        j.t_launch = 0;
        sv->pending.push_back(j);
        sv->submitted++;
    }
    sv->cv_work.notify_one();
    return RAV2D_OK;
}

int rav2d_gpu_server_poll(Rav2dGpuServer *sv, Rav2dServeResult *out, int max, int wait_ms) {
    if (!sv || (!out && max > 0) || max < 0) return RAV2D_EINVAL;
    std::unique_lock<std::mutex> lk(sv->mu);
    if (sv->done.empty() && wait_ms > 0)
        sv->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !sv->done.empty(); });
    int n = 0;
    while (n < max && !sv->done.empty()) {
        out[n++] = sv->done.front();
        sv->done.pop_front();
    }
    return n;
}

int rav2d_gpu_server_pause(Rav2dGpuServer *sv, int paused) {
    if (!sv) return RAV2D_EINVAL;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->paused = paused != 0;
    }
    sv->cv_work.notify_all();
    return RAV2D_OK;
}

int rav2d_gpu_server_cancel(Rav2dGpuServer *sv, uint64_t tag) {
    if (!sv) return RAV2D_EINVAL;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = std::find_if(sv->pending.begin(), sv->pending.end(), [&](const Job &j) { return j.tag == tag; });
        if (it == sv->pending.end()) {
            set_last_error("server_cancel: no queued job with this tag (launched, done or unknown)");
            return RAV2D_ENOENT;
        }
        Rav2dServeResult res;
        memset(&res, 0, sizeof(res));
        res.tag = it->tag;
        res.t_submit = it->t_submit;
        res.t_done = sv->now();
        res.status = RAV2D_ECANCELED;
        sv->pending.erase(it);
        sv->done.push_back(res);
        sv->completed++;
    }
    sv->cv_done.notify_all();
    return RAV2D_OK;
}

int rav2d_gpu_server_drain(Rav2dGpuServer *sv) {
    if (!sv) return RAV2D_EINVAL;
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv_done.wait(lk, [&] { return sv->completed == sv->submitted; });
    return RAV2D_OK;
}

int rav2d_gpu_server_simulate(Rav2dGpuServer *sv, Rav2dGpuFrame *const *frames, int n_streams, int slots,
                              double fps, int n_jobs, uint64_t seed, Rav2dServeResult *results,
                              int32_t *job_stream, int32_t *job_slot) {
    if (!sv || !frames || n_streams <= 0 || slots <= 0 || fps <= 0 || n_jobs <= 0 || !results) return RAV2D_EINVAL;
    for (int i = 0; i < n_streams * slots; i++)
        if (!frames[i] || frames[i]->ctx != sv->ctx) return RAV2D_EINVAL;
    // Poisson arrivals per stream, merged in time order
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(fps);
    typedef std::pair<double, int> Ev;
    std::priority_queue<Ev, std::vector<Ev>, std::greater<Ev>> q;
    for (int s = 0; s < n_streams; s++) q.push({gap(rng), s});
    std::vector<double> t(n_jobs);
    std::vector<int> st(n_jobs), sl(n_jobs), count(n_streams, 0);
    for (int j = 0; j < n_jobs; j++) {
        Ev e = q.top();
        q.pop();
        t[j] = e.first;
        st[j] = e.second;
        sl[j] = count[e.second]++ % slots;
        q.push({e.first + gap(rng), e.second});
    }
    const double start = sv->now() + 0.005;
    std::vector<Rav2dServeResult> buf(256);
    int got = 0;
    auto collect = [&](int wait_ms) {
        int n = rav2d_gpu_server_poll(sv, buf.data(), (int)buf.size(), wait_ms);
        for (int i = 0; i < n; i++) {
            uint64_t tag = buf[i].tag;
            if (tag < (uint64_t)n_jobs) results[tag] = buf[i];
        }
        got += std::max(n, 0);
    };
    for (int j = 0; j < n_jobs; j++) {
        const double due = start + t[j];
        for (;;) {
            double now = sv->now();
            if (now >= due) break;
            collect(0);
            double left = due - sv->now();
            if (left > 300e-6) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)((left - 200e-6) * 1e6)));
        }
        int r = rav2d_gpu_server_submit(sv, frames[st[j] * slots + sl[j]], (uint64_t)j, due);
        if (r) return r;
        if (job_stream) job_stream[j] = st[j];
        if (job_slot) job_slot[j] = sl[j];
    }
    while (got < n_jobs) collect(100);
    return RAV2D_OK;
}

int rav2d_gpu_frame_checksum(Rav2dGpuFrame *f, int which, uint64_t *out) {
    if (!f || !out || which < 0 || which > 2) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(f->ctx->device));
    hipStream_t s = f->ctx->stream;
    const int b = which == 0 ? BUF_IN : which == 1 ? BUF_OUT : BUF_DISP;
    uint8_t *mem = nullptr;
    const size_t bytes = 256 + sizeof(SumArgs) + 256;
    HIP_TRY(hipMalloc((void **)&mem, bytes));
    unsigned long long *d_out = (unsigned long long *)mem;
    int32_t *d_prefix = (int32_t *)(mem + 64);
    SumArgs *d_sum = (SumArgs *)(mem + 256);
    SumArgs a = sum_args(f, b, 1, d_out);
    int32_t zero = 0;
    unsigned long long h = 0;
    int rc = RAV2D_OK;
    if (hipMemcpyAsync(d_sum, &a, sizeof(a), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_prefix, &zero, sizeof(zero), hipMemcpyHostToDevice, s) != hipSuccess) {
        rc = RAV2D_EDEVICE;
    } else {
        hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, s, d_out, 2);
        BatchSched bs;
        bs.prefix = d_prefix;
        bs.n_frames = 1;
        bs.total = a.first[3];
        bs.xcd_remap = 0;
        hipLaunchKernelGGL(checksum_batch, dim3(a.first[3]), dim3(256), 0, s, d_sum, bs);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&h, d_out, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = RAV2D_EDEVICE;
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(mem);
    if (rc) {
        set_last_error("frame checksum failed");
        return rc;
    }
    *out = h;
    return which == 2 ? frame_errors(f) : RAV2D_OK;
}

}  // extern "C"
