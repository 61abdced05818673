// Device-side shared definitions for the gfx950 filter kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rav2d_gpu.h"

namespace r2g {

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

template <typename P> __device__ __forceinline__ P *row_ptr(const DevPlane &p, int y) {
    return (P *)(p.ptr + (int64_t)y * p.stride);
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

}  // namespace r2g
