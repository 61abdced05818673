// Host-side internals shared by the runtime and the kernel launchers.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "dev_common.h"

namespace r2g {

// Argument arrays of a mixed launch: frame argument blocks of each bit-depth
// group (8-, 10-, 12-bit) in device memory.
template <typename T> struct DevArgs3 {
    const T *p[3];
};

// Per-frame deblock threshold LUT: [kind][qidx] -> (q_thr, side_thr); kinds
// 0 = luma vertical edges, 1 = luma horizontal edges, 2 = U, 3 = V.
// init_deblock_thr_lut_y/uv (crates/rav2d/src/deblock.rs:13-57) evaluated for
// every qidx a 64x64 block (or, segmented, a segment of it) can carry.
struct DbLut {
    uint16_t q[4][512];
    uint16_t side[4][512];
};

struct DeblockArgs {
    DevPlane src[3], dst[3];
    DevGeom g;
    const Rav2dPackedFilter *mask;
    const ChromaMasks *uvm;  // 4:2:2 / 4:4:4: the chroma masks (null: the packed record's 4:2:0 ones)
    const DbLut *lut;
    int32_t plane_on[3];  // plane filtered in this pass
    TileMap map;
    // picture padding read below / right of every plane (rows, samples); the
    // vertical-edge pass also carries the rows below into its output, where
    // the horizontal-edge pass reads them
    int32_t pad_rows, pad_cols;
    // segmentation (null seg[0]: off): segment id per 4x4 unit of luma / chroma;
    // a unit's LUT row is clip(qidx + seg_dq[id], 0, qmax), as the per-segment
    // LUT of init_deblock_thr_lut_y/uv (deblock.rs:13-57) gives
    const uint8_t *seg[2];
    int32_t seg_stride[2];
    int32_t qmax;
    int16_t seg_dq[8];
};

// CCSO configuration of one plane (build_ccso_plane_cfg, decode.rs:4117-4133).
struct CcsoDev {
    int32_t shift;       // bitdepth - max_band_log2
    int32_t dx, dy;      // CCSO_POS[ext_filter_support]
    int32_t quant_step;  // CCSO_QUANT_SZ[scale_idx][quant_idx]
    uint32_t edge_clf;
    int32_t bo_only;
    alignas(4) int8_t lut[256];  // offset per band/edge index: CCSO_OFFSET[scale_idx] of filter_off's nibble
};

struct CdefArgs {
    DevPlane src[3], dst[3];
    DevGeom g;
    const Rav2dPackedFilter *mask;
    const ChromaMasks *uvm;  // 4:2:2 / 4:4:4: the chroma masks (null: the packed record's 4:2:0 ones)
    const int32_t *order;  // superblock of each launch slot (null: raster order)
    int32_t cdef_on, ccso_on, on_skiptx, damping;
    int32_t any_ll;  // a segment is lossless: CDEF and CCSO skip the masks' lossless 4x4 units
    int32_t has_chroma;  // 0: monochrome (I400), no chroma planes to stage or filter
    alignas(8) uint8_t y_strength[8];
    alignas(8) uint8_t uv_strength[8];
    CcsoDev ccso[3];
    alignas(4) int16_t off[2][12][2];  // CDEF_DIRECTIONS as offsets in the kernel's luma / chroma tiles
};

void launch_cdef(const CdefArgs &a, int bitdepth, hipStream_t s);
void launch_cdef_batch(const CdefArgs *d_args, const BatchSched &s, int grid, int bitdepth, bool any, hipStream_t st);
int cdef_items(const CdefArgs &a);  // 64x64 superblocks of the frame
// Host side of the argument block: CCSO's offset table of one plane (from the
// plane's packed filter_off nibbles and scale_idx), and the direction offsets.
void cdef_ccso_lut(CcsoDev &d, const uint8_t filter_off[64], int scale_idx);
void cdef_tap_offsets(CdefArgs &a);

// Loop restoration: one work item per (plane, stripe, 64-column unit), built
// on the host from the stripe/unit walk of lr_sbrow_apply / lr_sbrow_plane /
// lr_stripe (crates/rav2d/src/looprestoration.rs:1847-2548).
enum { LR_KIND_NONE = 0, LR_KIND_SINGLE = 1, LR_KIND_MULTI = 2, LR_KIND_PC = 3 };
enum { LRE_LEFT = 1, LRE_RIGHT = 2, LRE_TOP = 4, LRE_BOTTOM = 8 };

struct LrItem {
    int16_t plane, x0, w, y0, h;  // stripe rows [y0, y0+h) x unit columns [x0, x0+w)
    uint8_t edges, kind, gdf_on, pad;
    int8_t filt[18];              // single-filter taps (luma 16 / chroma 18)
    uint16_t noskip[17];          // lr_noskip rows of the stripe (multi-class kinds)
    uint16_t ll[17];              // lossless rows of the stripe, bit per 4x4 column
    // chroma NS-Wiener items of 4:2:2 / 4:4:4 frames: the deblocked luma rows the
    // reference's luma line store holds where the item reads it (above the
    // stripe, its wrap-around row, below it twice, that one's wrap-around row);
    // -1: a store row never written (zeros)
    int16_t lsrc[5];
    int16_t pad2;
};

// Device copies of the loop-restoration / GDF tables (tables.h).
struct LrTables {
    const uint8_t *lut_to_class;    // PC_WIENER_LUT_TO_CLASS[4096]
    const int16_t *pc_filters;      // PC_WIENER_FILTERS[4][64][13]
    const uint32_t *pc_pairs;       // the same as [4][64][8] packed 16-bit pairs (f0|f1, .., f10|f11, f12|0, 0)
    const uint8_t *cls_pc;          // [4][4096]: PC_WIENER_SUB_CLASSIFY[w][LUT_TO_CLASS[li]]
    const uint8_t *cls_ns;          // [4][7][4096]: PC_WIENER_SUB_CLASSIFY_NS[w][n][LUT_TO_CLASS[li]]
    const uint8_t *pc_sub;          // PC_WIENER_SUB_CLASSIFY[4][256]
    const uint8_t *ns_sub;          // PC_WIENER_SUB_CLASSIFY_NS[4][7][256]
    const uint16_t *gdf_alpha;      // GDF_ALPHA[3168]
    const int16_t *gdf_weight;      // GDF_WEIGHT[9504]
    const int16_t *gdf_bias;        // GDF_BIAS[36][3]
    const int8_t *gdf_intra;        // GDF_INTRA_ERROR[24576]
    const int8_t *gdf_inter;        // GDF_INTER_ERROR[30000]
};

struct LrArgs {
    DevPlane f3[3], f2[3], dst[3];
    DevGeom g;
    const LrItem *items;
    LrTables t;
    int32_t cfl_ds;
    int32_t base_q;
    int32_t widx, ns_idx;
    int32_t gdf_ref, gdf_qp, gdf_scale;
    int32_t lstore_stride;    // samples per row of the emulated luma line store (the luma plane's stride)
    int8_t mfilt[3][16][18];  // frame filters of each plane (multi-class NS)
};

void launch_lr(const LrArgs &a, int n_items, int bitdepth, hipStream_t s);
void launch_lr_batch(const LrArgs *d_args, const BatchSched &s, int grid, int bitdepth, hipStream_t st);
int lr_tables(int device, LrTables *out);

// Film grain (filmgrain.rs): LUTs, scaling and block offsets live in a
// per-frame device scratch area written by the setup kernel.
// Error flag of the set-up: a chroma AR wave's bounded wait for the luma
// wave's progress ran out (the luma wave never waits, so this is not expected).
constexpr uint32_t kGrainErrLumaWait = 1u;
struct GrainArgs {
    DevPlane src[3], dst[3];
    Rav2dFilmGrainData fg;
    uint32_t seed;
    int32_t bitdepth;
    int16_t *lut;        // [3][73][82]
    uint8_t *scaling;    // [3][1024]
    uint8_t *offsets;    // [n_bands][n_blocks][2]
    uint32_t *err;       // error flags the set-up raises (kGrainErr*); may be null
    int32_t n_bands, n_blocks;
    int32_t ss_hor, ss_ver;  // chroma subsampling
    int32_t plane_on[3];
    TileMap map;
};

void launch_grain_setup(const GrainArgs &a, hipStream_t s);
void launch_grain_apply(GrainArgs a, int bitdepth, hipStream_t s);
void launch_grain_setup_batch(const GrainArgs *d_args, int n, hipStream_t s);
void launch_grain_apply_batch(const GrainArgs *d_args, const BatchSched &s, int grid, int bitdepth, bool any,
                              hipStream_t st);
void grain_tile_map(GrainArgs &a);  // 64x32 tiles of the three planes
bool grain_init_tables();           // Gaussian table and jump matrices of the current device
// generate_grain_y (plane 0) / generate_grain_uv (plane 1 U, 2 V; 4:2:0, from
// the luma template `luma`) of one plane into `out` (rows of 82, device).
int launch_grain_generate(const Rav2dFilmGrainData &fg, uint32_t seed, int bitdepth, int plane, int ss_hor, int ss_ver,
                          int16_t *out, const int16_t *luma, hipStream_t s);

// Both deblocking passes in one tile pass: F0 -> F2 without the intermediate
// plane.  `a` carries the vertical-edge pass's arguments (source F0, masks,
// LUT, padding, segmentation), which the horizontal-edge pass shares; on_v /
// on_h say which planes each pass filters (a plane neither filters is
// copied).  The tiles are 64 x rows (32 or 64) of the output planes.
struct DeblockVHArgs {
    DeblockArgs a;
    DevPlane dst[3];
    int32_t on_v[3], on_h[3];
    TileMap map;
};
void launch_deblock_v(const DeblockArgs &a, int bitdepth, hipStream_t s);
void launch_deblock_h(const DeblockArgs &a, int bitdepth, hipStream_t s);
// Tile rows the fused pass supports.
inline bool deblock_vh_rows_ok(int rows) { return rows == 32 || rows == 64; }
TileMap deblock_vh_tile_map(const DeblockArgs &a, int rows);
void launch_deblock_vh(const DeblockVHArgs &a, int bitdepth, int rows, hipStream_t s);
// `any`: some frame of the launch has 4:2:2 / 4:4:4 chroma (the general kernels)
void launch_deblock_vh_batch(const DeblockVHArgs *d_args, const BatchSched &s, int grid, int bitdepth, int rows,
                             bool any, hipStream_t st);
// One launch over the frames of every bit-depth group (MixSched).
void launch_deblock_mix(bool vert, DevArgs3<DeblockArgs> args, const MixSched &s, int grid, bool any, hipStream_t st);
void launch_deblock_vh_mix(DevArgs3<DeblockVHArgs> args, const MixSched &s, int grid, int rows, bool any,
                           hipStream_t st);
void launch_cdef_mix(DevArgs3<CdefArgs> args, const MixSched &s, int grid, bool any, hipStream_t st);
void launch_lr_mix(DevArgs3<LrArgs> args, const MixSched &s, int grid, hipStream_t st);
void launch_grain_apply_mix(DevArgs3<GrainArgs> args, const MixSched &s, int grid, bool any, hipStream_t st);
void launch_deblock_batch(bool vert, const DeblockArgs *d_args, const BatchSched &s, int grid, int bitdepth,
                          bool any, hipStream_t st);
// 64x32 tiles of the three planes (the vertical-edge pass: down to the padding rows)
TileMap deblock_tile_map(const DeblockArgs &a, bool vert);

void set_last_error(const std::string &msg);
// rav2d_gpu_debug_inject_failure: true when the next operation of kind `what`
// (RAV2D_FAIL_ALLOC / RAV2D_FAIL_LAUNCH) is to fail
bool injected_failure(int what);

// Kernel-accurate timing of one launch.  While armed, the next kernel launched
// through R2G_LAUNCH on this thread carries the two events, which the runtime
// stamps at that kernel's own start and end (hipExtLaunchKernelGGL), so the
// interval excludes launch gaps; the slot is then disarmed.
struct LaunchTiming {
    hipEvent_t start = nullptr, stop = nullptr;
    bool used = false;
};
LaunchTiming &launch_timing();

#define R2G_LAUNCH(K, G, B, SH, S, ...)                                                    \
    do {                                                                                   \
        ::r2g::LaunchTiming &lt_ = ::r2g::launch_timing();                                 \
        if (lt_.start) {                                                                   \
            hipEvent_t a_ = lt_.start, b_ = lt_.stop;                                      \
            lt_.start = lt_.stop = nullptr;                                                \
            lt_.used = true;                                                               \
            hipExtLaunchKernelGGL(K, G, B, SH, S, a_, b_, 0u, __VA_ARGS__);                \
        } else {                                                                           \
            hipLaunchKernelGGL(K, G, B, SH, S, __VA_ARGS__);                               \
        }                                                                                  \
    } while (0)

}  // namespace r2g
