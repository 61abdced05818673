/*
 * Seeded synthetic filter-stage workloads: reconstructed planes plus the side
 * info a decoder would hand to the filter stage (Av2Filter / Av2Restoration
 * records and frame parameters).  Used by the tests and the benchmark; not
 * part of the drop-in ABI in rav2d_gpu.h.
 *
 * Planes: smooth 2-D gradient + Gaussian noise (sigma 4 at 8 bit, 16 at 10, 64 at 12 bit)
 * + per-transform-block DC offsets on a random quadtree of 8x8..64x64 blocks with
 * square transforms 4..64.  Edge masks are produced the way create_db_mask
 * (crates/rav2d/src/lf_mask.rs:399-502) does: by default for intra blocks
 * without transform partitioning; with R2SConfig.mask_flags also for the
 * transform partitions (mask_edges_part, :156-254) and the sub-PU edges of
 * inter blocks (mask_subpu_edges, :256-330); parameters are drawn from the legal ranges of the frame header
 * parser (crates/rav2d/src/obu.rs:261-328, 1293-1455, 1696-1736).
 */
#ifndef RAV2D_SYNTH_H
#define RAV2D_SYNTH_H
#include "../../include/rav2d_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* which tools the generated frame enables */
#define R2S_TOOL_DEBLOCK (1u << 0)
#define R2S_TOOL_CDEF (1u << 1)
#define R2S_TOOL_CCSO (1u << 2)
#define R2S_TOOL_WIENER (1u << 3)
#define R2S_TOOL_GDF (1u << 4)
#define R2S_TOOL_GRAIN (1u << 5)
#define R2S_TOOL_ALL 63u

/* mask shapes beyond the default square-transform intra blocks */
/* transform partitions of TX_PART_TBL (None, Split, H, V, H4, V4, H5, V5) */
#define R2S_MASK_TXPART (1u << 0)
/* inter blocks with sub-PU edges (subpu_flt_lvl 0..2) and skip blocks that
 * carry only outer and sub-PU edges; a quarter of such frames are TIP frames
 * of frame_mode 2 (no sub-PU threshold bit, ds_sub_pu_mask = 0) */
#define R2S_MASK_SUBPU (1u << 1)
/* segmentation: random per-segment qidx deltas, segment ids per block in the
 * segment maps; lossless blocks (lossless_pct) form a lossless segment 7 */
#define R2S_MASK_SEG (1u << 2)

typedef struct R2SConfig {
    int32_t width;
    int32_t height;
    int32_t bitdepth;
    int32_t sb128;
    uint64_t seed;
    uint32_t tools;          /* R2S_TOOL_* */
    int32_t lossless_pct;    /* percent of blocks whose masks are marked lossless */
    int32_t skip_pct;        /* percent of blocks without residual (noskip = 0)   */
    int32_t pc_gdf_pct;      /* percent of restoration units / GDF blocks enabled */
    uint32_t mask_flags;     /* R2S_MASK_*; 0 = square transforms only            */
    int32_t monochrome;      /* 1: I400 (luma only; no chroma planes, chroma tools
                                off, as a monochrome stream's headers give); 0: chroma_format */
    int32_t chroma_format;   /* 0: 4:2:0, 1: 4:2:2, 2: 4:4:4.  Every draw is the same in
                                each format, so the luma planes, luma side info and luma
                                parameters equal those of the 4:2:0 frame of the seed */
} R2SConfig;

typedef struct R2SFrame R2SFrame;

R2SFrame *r2s_frame_create(const R2SConfig *cfg);
void r2s_frame_destroy(R2SFrame *f);
Rav2dFrameParams *r2s_params(R2SFrame *f);
/* tightly packed planes: stride = plane width in samples */
void *r2s_plane(R2SFrame *f, int pl);
/* the same plane inside a picture buffer with 8 rows below and 8 columns
 * right of it (the padding of a decoder's picture, which holds the
 * reconstruction of blocks that overhang the frame and which the deblocking
 * reads); *stride = (width + 8) samples in bytes */
void *r2s_plane_padded(R2SFrame *f, int pl, ptrdiff_t *stride);
int32_t r2s_plane_width(const R2SFrame *f, int pl);
int32_t r2s_plane_height(const R2SFrame *f, int pl);
size_t r2s_plane_bytes(const R2SFrame *f, int pl);
Rav2dAv2Filter *r2s_mask(R2SFrame *f, size_t *n);
Rav2dAv2Restoration *r2s_lr(R2SFrame *f, size_t *n);
/* segment map of a segmented frame (which 0: luma 4x4 units, 1: chroma), NULL
 * when segmentation is off; *stride = bytes per row */
uint8_t *r2s_segmap(R2SFrame *f, int which, ptrdiff_t *stride);

/* The generator's mask producers on one [2][64][5][4] mask set (the
 * FilterMasks pair of lf_mask.rs), for unit tests: mask_edges_part
 * (lf_mask.rs:156-254; part = TxPartition 0..7, tw4/th4 the transform size in
 * 4px units, a / l the above / left level contexts) and mask_subpu_edges
 * (lf_mask.rs:256-330). */
void r2s_mask_edges_part(uint16_t (*masks)[64][5][4], int by4, int bx4, int w4, int h4, int part, int tw4, int th4,
                         int hlim, int vlim, uint8_t *a, uint8_t *l);
void r2s_mask_subpu_edges(uint16_t (*masks)[64][5][4], int by4, int bx4, int w4, int h4, int twl4c, int thl4c,
                          int hsz, int vsz, int ds_mask);

#ifdef __cplusplus
}
#endif
#endif
