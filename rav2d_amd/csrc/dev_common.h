// Device-side shared definitions for the gfx950 filter kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rav2d_gpu.h"

namespace r2g {

// The chroma edge and lossless masks of one SB256 at full size, the device
// copy of a 4:2:2 or 4:4:4 frame's Av2Filter chroma fields (lf_mask.rs:9, 18):
// lines and rows are the plane's 4x4 units in the SB256 (64 >> subsampling),
// filter_uv words hold 16 units each, lossless words one 64-luma-column band.
struct ChromaMasks {
    uint16_t filter_uv[2][64][5][4];
    uint16_t lossless_mask_uv[64][4];
};

// Chroma subsampling of a layout (I400 and I420: both; I422: horizontal).
__host__ __device__ inline int layout_ss_hor(int layout) { return layout != RAV2D_PIXEL_LAYOUT_I444; }
__host__ __device__ inline int layout_ss_ver(int layout) {
    return layout == RAV2D_PIXEL_LAYOUT_I420 || layout == RAV2D_PIXEL_LAYOUT_I400;
}

// One plane of one frame as the kernels see it.
struct DevPlane {
    uint8_t *ptr;     // first sample
    int64_t stride;   // bytes
    int32_t w, h;     // plane size in samples (filtered region)
};

// Frame geometry shared by all passes (decode.rs:4244-4251).
struct DevGeom {
    int32_t bw, bh;          // luma size in 4px units (rounded up to 8 px)
    int32_t sb256w, sb256h;  // SB256 grid
    int32_t sb128;
    int32_t bitdepth, bdmax, hbd;
    int32_t ss_hor, ss_ver;
};

template <typename T> __device__ __forceinline__ T dmin(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T dmax(T a, T b) { return a > b ? a : b; }
__device__ __forceinline__ int iclip(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int apply_sign(int v, int s) { return s < 0 ? -v : v; }
__device__ __forceinline__ int ulog2(uint32_t v) { return 31 ^ __clz((int)v); }

template <typename P, typename D> __device__ __forceinline__ P *row_ptr(const D &p, int y) {
    return (P *)(p.ptr + (int64_t)y * p.stride);
}

// Four consecutive samples (4-sample aligned) <-> four u16 packed in a uint2,
// the LDS tile layout.
template <typename P> __device__ __forceinline__ uint2 load4(const P *p) {
    if constexpr (sizeof(P) == 1) {
        const uint32_t b = *(const uint32_t *)p;
        return make_uint2((b & 0xff) | ((b << 8) & 0xff0000), ((b >> 16) & 0xff) | ((b >> 8) & 0xff0000));
    } else {
        return *(const uint2 *)p;
    }
}
template <typename P> __device__ __forceinline__ void store4(P *p, uint2 v) {
    if constexpr (sizeof(P) == 1)
        *(uint32_t *)p = (v.x & 0xff) | ((v.x >> 8) & 0xff00) | ((v.y & 0xff) << 16) | ((v.y << 8) & 0xff000000);
    else
        *(uint2 *)p = v;
}

// One 16-bit LDS read that the compiler keeps separate from its neighbours (a
// relaxed workgroup-scope atomic load: not merged into a possibly unaligned
// wider access, and still a plain ds_read with no waits of its own).  `p`
// must point into LDS.
__device__ __forceinline__ int16_t lds_load_i16(const int16_t *p) {
    return __hip_atomic_load((__attribute__((address_space(3))) int16_t *)p, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Workgroup b of a one-dimensional launch runs on XCD b mod 8: the dispatcher
// deals workgroups round-robin over the 8 XCDs, each with its own L2.  This
// maps b to a work item so that each XCD takes a contiguous run of items and
// neighbouring tiles, whose halos overlap, share an L2: the first 8 * (n / 8)
// workgroups as XCD-major runs, the remaining n mod 8 unchanged (a bijection
// of [0, n)).
__device__ __forceinline__ int xcd_item(int b, int n) {
    const int q = n >> 3;
    return b < (q << 3) ? (b & 7) * q + (b >> 3) : b;
}

// Two adjacent samples (2-sample aligned) with one store.
template <typename P> __device__ __forceinline__ void store2(P *p, int a, int b) {
    if constexpr (sizeof(P) == 1)
        *(uint16_t *)p = (uint16_t)((a & 0xff) | ((b & 0xff) << 8));
    else
        *(uint32_t *)p = (uint32_t)(a & 0xffff) | ((uint32_t)b << 16);
}

// Argument blocks of batched launches live in device memory that no kernel
// writes; reading them through the constant address space lets the compiler
// use scalar loads (and reload instead of holding fields in registers).
template <typename T> using ConstArg = const __attribute__((address_space(4))) T;
template <typename T> __device__ __forceinline__ ConstArg<T> &const_arg(const T *base, int i) {
    return *(ConstArg<T> *)(base + i);
}

// Tile schedule over the three planes of one frame: tiles of plane 0 first.
struct TileMap {
    int32_t tiles_x[3], tiles_y[3];
    int32_t first[4];  // prefix of tile counts; first[3] = total
};

__device__ __forceinline__ void decode_tile(const TileMap &m, int t, int &pl, int &tx, int &ty) {
    pl = t >= m.first[2] ? 2 : (t >= m.first[1] ? 1 : 0);
    int r = t - m.first[pl];
    ty = r / m.tiles_x[pl];
    tx = r - ty * m.tiles_x[pl];
}

// Work schedule of a batched launch: work items of all frames back to back;
// prefix[f] is frame f's first item.  A launch may be persistent (fewer
// workgroups than items, each looping); with xcd_remap the items are cut into
// 8 contiguous chunks, one per XCD (workgroup b runs on XCD b % 8), so the
// neighbouring tiles that share halo rows share an L2.
struct BatchSched {
    const int32_t *prefix;
    int32_t n_frames;
    int32_t total;
    int32_t xcd_remap;
};

// Frame of a batch item: prefix is increasing with prefix[0] = 0, so the frame
// is the number of entries <= item, minus one.  Each lane of the wave reads one
// entry (one round of loads for up to 64 frames, instead of a binary search's
// chain of dependent loads) and a ballot counts them.  Called by whole waves
// with a wave-uniform item.
__device__ __forceinline__ int find_frame(const int32_t *prefix, int n, int item) {
    const int lane = threadIdx.x & 63;
    int count = 0;
    for (int base = 0; base < n; base += 64) {
        const int i = base + lane;
        const bool le = i < n && prefix[i] <= item;
        count += __popcll(__ballot(le));
    }
    return count - 1;
}

// Visits the items of workgroup blockIdx.x in order: fn(frame, local item).
// kLoop = false: one item per workgroup (the grid covers every item);
// kLoop = true: a persistent grid, each workgroup loops.  Frame and item are
// workgroup-uniform; readfirstlane tells the compiler so.
template <bool kLoop, typename F> __device__ __forceinline__ void for_each_item(const BatchSched &s, F fn) {
    const int G = gridDim.x, b = blockIdx.x;
    int hi = s.total, step = G, first = b;
    if (s.xcd_remap && (G & 7) == 0) {
        const int per_xcd = G >> 3, xcd = b & 7, slot = b >> 3;
        const int chunk = (s.total + 7) >> 3;
        first = xcd * chunk + slot;
        hi = min(s.total, xcd * chunk + chunk);
        step = per_xcd;
    }
    if (!kLoop) {
        if (first < hi) {
            const int f = __builtin_amdgcn_readfirstlane(find_frame(s.prefix, s.n_frames, first));
            fn(f, __builtin_amdgcn_readfirstlane(first - s.prefix[f]));
        }
        return;
    }
    for (int it = first; it < hi; it += step) {
        const int f = __builtin_amdgcn_readfirstlane(find_frame(s.prefix, s.n_frames, it));
        fn(f, __builtin_amdgcn_readfirstlane(it - s.prefix[f]));
    }
}

// Schedule of one launch over the frames of every bit-depth group (8-, 10-
// and 12-bit): group g's items follow those of the groups before it, from
// item first[g]; g[i] is group i's own schedule (prefix of its frames).
struct MixSched {
    BatchSched g[3];
    int32_t first[4];  // first[3] = total
    int32_t xcd_remap;
};

// Visits the items of workgroup blockIdx.x of a mixed launch in order:
// fn(group, frame, local item), all workgroup-uniform.
template <bool kLoop, typename F> __device__ __forceinline__ void for_each_mixed_item(const MixSched &s, F fn) {
    const int G = gridDim.x, b = blockIdx.x, total = s.first[3];
    int hi = total, step = G, first = b;
    if (s.xcd_remap && (G & 7) == 0) {
        const int per_xcd = G >> 3, xcd = b & 7, slot = b >> 3;
        const int chunk = (total + 7) >> 3;
        first = xcd * chunk + slot;
        hi = min(total, xcd * chunk + chunk);
        step = per_xcd;
    }
    for (int it = first; it < hi; it += step) {
        const int g = __builtin_amdgcn_readfirstlane(it >= s.first[2] ? 2 : (it >= s.first[1] ? 1 : 0));
        const int li = it - s.first[g];
        const int f = __builtin_amdgcn_readfirstlane(find_frame(s.g[g].prefix, s.g[g].n_frames, li));
        fn(g, f, __builtin_amdgcn_readfirstlane(li - s.g[g].prefix[f]));
        if (!kLoop) break;
    }
}

}  // namespace r2g
