// HBM bandwidth probe: a plain streaming copy of known size, timed with HIP
// events.  It gives the achievable copy bandwidth next to the 8 TB/s
// datasheet peak, and a reference kernel whose HBM counters (FETCH_SIZE /
// WRITE_SIZE) scale the filter kernels' counters to bytes.
#include "frame.h"

namespace {

__global__ __launch_bounds__(256) void hbm_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                      size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t step = (size_t)gridDim.x * 256;
    for (; i < n; i += step) dst[i] = src[i];
}

}  // namespace

using namespace r2g;

extern "C" int rav2d_gpu_bandwidth_probe(Rav2dGpuCtx *ctx, uint64_t bytes, int iters, double *gbps) {
    if (!ctx || !gbps || iters <= 0 || bytes < 4096 || (bytes & 15)) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void **)&a, bytes) != hipSuccess || hipMalloc((void **)&b, bytes) != hipSuccess) {
        if (a) (void)hipFree(a);
        set_last_error("bandwidth probe: hipMalloc failed");
        return RAV2D_ENOMEM;
    }
    const size_t n = bytes / 16;
    const int grid = 256 * 16;
    hipStream_t s = ctx->stream;
    int rc = RAV2D_OK;
    (void)hipMemsetAsync(a, 1, bytes, s);
    hipLaunchKernelGGL(hbm_copy_probe, dim3(grid), dim3(256), 0, s, a, b, n);  // warm
    (void)hipEventRecord(ctx->ev0, s);
    for (int i = 0; i < iters; i++) hipLaunchKernelGGL(hbm_copy_probe, dim3(grid), dim3(256), 0, s, a, b, n);
    (void)hipEventRecord(ctx->ev1, s);
    float ms = 0.f;
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(ctx->ev1) != hipSuccess ||
        hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) {
        set_last_error("bandwidth probe failed");
        rc = RAV2D_EDEVICE;
    } else {
        *gbps = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;  // read + write
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}
