// Host-side internals shared by the runtime and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "dev_common.h"

namespace r2g {

// Per-frame deblock threshold LUT: [kind][qidx] -> (q_thr, side_thr); kinds
// 0 = luma vertical edges, 1 = luma horizontal edges, 2 = U, 3 = V.
// init_deblock_thr_lut_y/uv (crates/rav2d/src/deblock.rs:13-57) with
// segmentation disabled, evaluated for every qidx a 64x64 block can carry.
struct DbLut {
    uint16_t q[4][512];
    uint16_t side[4][512];
};

struct DeblockArgs {
    DevPlane src[3], dst[3];
    DevGeom g;
    const Rav2dAv2Filter *mask;
    const DbLut *lut;
    int32_t plane_on[3];  // plane filtered in this pass
    TileMap map;
};

// CCSO configuration of one plane (build_ccso_plane_cfg, decode.rs:4117-4133).
struct CcsoDev {
    int32_t shift;       // bitdepth - max_band_log2
    int32_t dx, dy;      // CCSO_POS[ext_filter_support]
    int32_t quant_step;  // CCSO_QUANT_SZ[scale_idx][quant_idx]
    uint32_t edge_clf;
    int32_t bo_only;
    int32_t scale_idx;
    uint8_t filter_off[64];
};

struct CdefArgs {
    DevPlane src[3], dst[3];
    DevGeom g;
    const Rav2dAv2Filter *mask;
    int32_t cdef_on, ccso_on, on_skiptx, damping;
    uint8_t y_strength[8], uv_strength[8];
    CcsoDev ccso[3];
};

void launch_cdef(const CdefArgs &a, int bitdepth, hipStream_t s);

// Which buffer a pass reads and writes, per plane.
void launch_deblock_v(const DeblockArgs &a, int bitdepth, hipStream_t s);
void launch_deblock_h(const DeblockArgs &a, int bitdepth, hipStream_t s);

void set_last_error(const std::string &msg);

}  // namespace r2g
