/*
 * rav2d_gpu.h — C ABI of the MI355X (gfx950) post-reconstruction filter stage.
 *
 * The library replaces rav2d's filter stage, which today runs on the host:
 *   - the per-superblock-row filter loop of decode_frame_main
 *     (crates/rav2d/src/decode.rs:3597-3625, body filter_sbrow decode.rs:3718-4058:
 *     deblock cols -> deblock rows -> copy_db -> CDEF(+CCSO) -> loop restoration);
 *   - film-grain synthesis on the output copy, apply_grain_to_picture_mt
 *     (crates/rav2d/src/decode.rs:4672, called from Decoder::output_image
 *     crates/rav2d/src/decoder.rs:430-435).
 * The side-info structs below keep the memory layout of the reference's Rust types
 * field for field (a #[repr(C)] mirror on the Rust side binds them unchanged), so
 * the masks the host decoder already builds are handed over as they are.
 *
 * Conventions: every entry point returns 0 on success or a negative RAV2D_E*
 * code and never aborts.  Plane strides are in BYTES, as in rav2d's
 * PictureAllocation (crates/rav2d/src/picture.rs:43-48).  Streams are hipStream_t
 * passed as void*; NULL means the context's own stream.
 */
#ifndef RAV2D_GPU_H
#define RAV2D_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define RAV2D_OK 0
#define RAV2D_EINVAL (-22)  /* invalid or unsupported parameter            */
#define RAV2D_ENOMEM (-12)  /* device or host allocation failed            */
#define RAV2D_EDEVICE (-5)  /* a HIP call failed                           */
#define RAV2D_EAGAIN (-11)  /* admission rejected (memory budget / slots)  */
#define RAV2D_ENOENT (-2)   /* unknown ticket / stream                     */
#define RAV2D_ECANCELED (-125) /* the job was cancelled before it launched */

/* Largest frame width and height the library accepts (luma samples). */
#define RAV2D_MAX_FRAME_DIM 16384

/* ---- enums with the reference's values --------------------------------- */
/* PixelLayout, crates/rav2d/src/headers.rs:112-118 */
#define RAV2D_PIXEL_LAYOUT_I400 0
#define RAV2D_PIXEL_LAYOUT_I420 1
#define RAV2D_PIXEL_LAYOUT_I422 2
#define RAV2D_PIXEL_LAYOUT_I444 3

/* InloopFilterType bits, crates/rav2d/src/decoder.rs:22-41 and
 * crates/rav2d/src/looprestoration.rs:1697-1706 */
#define RAV2D_INLOOPFILTER_DEBLOCK (1u << 0)
#define RAV2D_INLOOPFILTER_CDEF (1u << 1)
#define RAV2D_INLOOPFILTER_CCSO (1u << 2)
#define RAV2D_INLOOPFILTER_WIENER (1u << 3)
#define RAV2D_INLOOPFILTER_GDF (1u << 4)
#define RAV2D_INLOOPFILTER_ALL 31u

/* RestorationType, crates/rav2d/src/looprestoration.rs:1671-1676 */
#define RAV2D_RESTORATION_NONE 0
#define RAV2D_RESTORATION_PC_WIENER 1
#define RAV2D_RESTORATION_NS_WIENER 2
#define RAV2D_RESTORATION_SWITCHABLE 3

/* ---- side info: layout-identical mirrors of the reference's types -------- */

/* Av2Filter, crates/rav2d/src/lf_mask.rs:7-19: one per 256x256 luma region,
 * row-major over the frame's SB256 grid.  filter_y/uv[dir][x4|y4][level 0..3,
 * 4 = sub-PU][sb64 index] are the deblock edge masks; the rest is CDEF, CCSO,
 * GDF and LR side info. */
typedef struct Rav2dAv2Filter {
    uint16_t filter_y[2][64][5][4];
    uint16_t filter_uv[2][64][5][4];
    uint16_t qidx[16];
    uint8_t gdf[16];
    int8_t cdef_idx[16];
    uint8_t ccso[3];
    uint8_t _pad0;
    uint16_t noskip_mask[32][4];
    uint16_t lr_noskip_mask[64][4];
    uint16_t lossless_mask_y[64][4];
    uint16_t lossless_mask_uv[64][4];
} Rav2dAv2Filter;

/* Device form of one Av2Filter record for 4:2:0 frames (8004 B against the
 * record's 12100 B): the same fields with the chroma masks cut to the SB256's
 * 32 x 32 chroma 4x4 units (filter_uv lines 0..31, words 0..1; lossless_mask_uv
 * rows 0..31), which is every unit a 4:2:0 producer writes
 * (create_db_mask, lf_mask.rs:399-502: chroma by4/bx4 are luma units >> 1).
 * The library keeps and uploads side info in this form; a decoder may write it
 * directly from its mask producers and hand it to
 * rav2d_gpu_frame_set_side_info_packed, or convert Av2Filter records with
 * rav2d_gpu_pack_filter.  The Av2Restoration records stay on the host (the
 * library builds its loop-restoration work list from them). */
typedef struct Rav2dPackedFilter {
    uint16_t filter_y[2][64][5][4];
    uint16_t filter_uv[2][32][5][2];
    uint16_t lossless_mask_y[64][4];
    uint16_t lossless_mask_uv[32][4];
    uint16_t noskip_mask[32][4];
    uint16_t lr_noskip_mask[64][4];
    uint16_t qidx[16];
    int8_t cdef_idx[16];
    uint8_t ccso[3];
    uint8_t _pad0;
    uint8_t gdf[16];
} Rav2dPackedFilter;

/* Av2RestorationUnit / Av2Restoration, crates/rav2d/src/lf_mask.rs:1-5, 38-41 */
typedef struct Rav2dRestorationUnit {
    uint8_t restoration_type;
    int8_t ns_filter[16][32];
} Rav2dRestorationUnit;

typedef struct Rav2dAv2Restoration {
    Rav2dRestorationUnit lr[3][16];
} Rav2dAv2Restoration;

/* FilmGrainData, crates/rav2d/src/headers.rs:509-525 */
typedef struct Rav2dFilmGrainData {
    int32_t chroma_scaling_from_luma;
    int32_t num_points[3];
    uint8_t points[3][14][2];
    int32_t scaling_shift;
    int32_t ar_coeff_lag;
    int8_t ar_coeffs[3][28];
    uint64_t ar_coeff_shift;
    int32_t grain_scale_shift;
    int32_t uv_mult[2];
    int32_t uv_luma_mult[2];
    int32_t uv_offset[2];
    int32_t overlap_flag;
    int32_t clip_to_restricted_range;
    int32_t mc_identity;
    int32_t block_size;
} Rav2dFilmGrainData;

/* FhCcsoPlane, crates/rav2d/src/headers.rs:650-662 (fields the filter reads) */
typedef struct Rav2dCcsoPlane {
    uint8_t enabled;
    uint8_t bo_only;
    uint8_t scale_idx;
    uint8_t quant_idx;
    uint8_t ext_filter_support;
    uint8_t edge_clf;
    uint8_t max_band_log2;
    uint8_t _pad;
    uint8_t filter_off[64];
} Rav2dCcsoPlane;

/* FhRestorationPlane + NSWienerPlane, crates/rav2d/src/headers.rs:628-643 */
typedef struct Rav2dRestorationPlane {
    uint8_t restoration_type;
    uint8_t frame_filters_on;
    uint8_t num_classes_idx;
    uint8_t num_classes;
    int8_t filter[16][18];
} Rav2dRestorationPlane;

/* The frame-header and sequence-header fields the filter stage reads
 * (FilterFrameParams decode.rs:3645-3709, LrContext looprestoration.rs:1785-1818,
 * FhDeblock/FhCdef/FhCcso/FhRestoration/FhGdf headers.rs:589-690). */
typedef struct Rav2dFrameParams {
    int32_t width;           /* frame width in luma samples            */
    int32_t height;          /* frame height in luma samples           */
    int32_t bitdepth;        /* 8, 10 or 12                             */
    int32_t layout;          /* RAV2D_PIXEL_LAYOUT_*: I420, I422, I444, or
                                I400 (luma only: no chroma planes,
                                pointers or strides needed; chroma side
                                info and chroma parameters are ignored)  */
    int32_t sb128;           /* superblock size 128 (1) or 64 (0)       */
    uint32_t inloop_filters; /* RAV2D_INLOOPFILTER_* mask               */
    /* sequence header */
    int32_t seq_cdef;             /* SequenceHeader.cdef                */
    int32_t cfl_ds_filter_index;  /* 0 box 2x2, 1 vertical avg, 2 copy  */
    /* quantizer: FhQuant */
    int32_t yac;
    int32_t uac_delta;
    int32_t vac_delta;
    /* FhSegmentation (headers.rs, parsed at obu.rs:1263-1280): with seg_enabled
     * the deblock thresholds of a 4x4 unit use its segment's qidx
     * clip(qidx + seg_delta_q[id], 0, 255 + 48 * hbd) (init_deblock_thr_lut_y/uv,
     * deblock.rs:13-57), from the segment maps of rav2d_gpu_frame_set_segmap;
     * CDEF and CCSO skip the lossless 4x4 units of the masks when a segment is
     * lossless (decode.rs:3868-3869, cdef.rs:894-926, 1117-1150). */
    int32_t seg_enabled;
    int16_t seg_delta_q[8];
    uint8_t seg_lossless[8];
    /* FhDeblock */
    uint8_t db_level_y[2];
    uint8_t db_level_u;
    uint8_t db_level_v;
    int8_t db_delta_q_y[2];
    int8_t db_delta_q_u;
    int8_t db_delta_q_v;
    /* FhCdef */
    uint8_t cdef_damping;
    uint8_t cdef_n_strengths;
    uint8_t cdef_on_skiptx;
    uint8_t _pad0;
    uint8_t cdef_y_strength[8];
    uint8_t cdef_uv_strength[8];
    /* FhCcso */
    Rav2dCcsoPlane ccso[3];
    /* FhRestoration */
    Rav2dRestorationPlane restoration[3];
    uint8_t lr_unit_size[2]; /* log2, luma / chroma                     */
    uint8_t _pad1[2];
    /* FhGdf + derived index (compute_gdf_ref_dst_idx decode.rs:67-81) */
    int32_t gdf_enabled;
    int32_t gdf_qp_idx;
    int32_t gdf_scale;
    int32_t gdf_ref_dst_idx;
    /* film grain: FhFilmGrain.present/seed + FilmGrainData */
    int32_t grain_present;
    uint32_t grain_seed;
    Rav2dFilmGrainData grain;
} Rav2dFrameParams;

/* Plane set in the PictureAllocation convention: 3 plane pointers, byte strides
 * (luma, chroma). */
typedef struct Rav2dPlanes {
    void *data[3];
    ptrdiff_t stride[2];
} Rav2dPlanes;

/* ---- context ------------------------------------------------------------ */
typedef struct Rav2dGpuCtx Rav2dGpuCtx;
typedef struct Rav2dGpuFrame Rav2dGpuFrame;

/* Number of HIP devices visible (0 when none or the runtime is absent). */
int rav2d_gpu_device_count(void);

/* Create a context on HIP device `device`.  hbm_budget_bytes bounds what the
 * frame pool and the serving admission may hold (0 = no bound). */
int rav2d_gpu_ctx_create(int device, uint64_t hbm_budget_bytes, uint32_t flags,
                         Rav2dGpuCtx **out);
int rav2d_gpu_ctx_destroy(Rav2dGpuCtx *ctx);
int rav2d_gpu_ctx_synchronize(Rav2dGpuCtx *ctx);
/* Bytes currently held by frames of this context. */
uint64_t rav2d_gpu_ctx_bytes_in_use(const Rav2dGpuCtx *ctx);

/* ---- device memory ---------------------------------------------------------
 * Device buffers counted against the context's HBM budget (RAV2D_EAGAIN when
 * it cannot hold them): for the per-family entries' buffers and for pictures
 * a decoder keeps in HBM.  rav2d_gpu_copy copies between host and device
 * memory in either direction (hipMemcpyDefault), enqueued on `stream`; with
 * stream NULL it runs on the context stream and returns once done. */
int rav2d_gpu_alloc(Rav2dGpuCtx *ctx, size_t bytes, void **out);
int rav2d_gpu_free(Rav2dGpuCtx *ctx, void *p);
int rav2d_gpu_copy(Rav2dGpuCtx *ctx, void *dst, const void *src, size_t bytes, void *stream);

/* ---- device-resident frames ---------------------------------------------
 * A frame owns: the reconstructed input planes (F0), the filtered output
 * planes (in-loop output, F4), optional display planes for the grained copy,
 * intermediate planes, and device copies of the side info.  It plays the part
 * of the PicAllocator pool of crates/rav2d/src/picture.rs:38-41, 77-125, with
 * its own layout: rows padded to 256 B (+256 B when the stride is a multiple
 * of 1 KiB), 8 spare rows per plane. */
int rav2d_gpu_frame_create(Rav2dGpuCtx *ctx, const Rav2dFrameParams *params,
                           Rav2dGpuFrame **out);
int rav2d_gpu_frame_destroy(Rav2dGpuFrame *f);
/* Upload side info: n_mask = sb256 rows * sb256 cols Av2Filter records,
 * n_lr = the same number of Av2Restoration records (lr may be NULL when no
 * plane uses Wiener restoration).  The records are packed on the host
 * (rav2d_gpu_pack_filter) and only the packed form goes to the device; for
 * 4:2:2 and 4:4:4 frames the chroma edge and lossless masks go to the device
 * at full size beside it. */
int rav2d_gpu_frame_set_side_info(Rav2dGpuFrame *f, const Rav2dAv2Filter *mask,
                                  size_t n_mask, const Rav2dAv2Restoration *lr,
                                  size_t n_lr, void *stream);
/* The same from records already in the device form (no repacking): 4:2:0 and
 * I400 frames (the packed form holds 4:2:0 chroma masks; RAV2D_EINVAL for
 * 4:2:2 and 4:4:4 frames, which take rav2d_gpu_frame_set_side_info). */
int rav2d_gpu_frame_set_side_info_packed(Rav2dGpuFrame *f, const Rav2dPackedFilter *mask,
                                         size_t n_mask, const Rav2dAv2Restoration *lr,
                                         size_t n_lr, void *stream);
/* Av2Filter records of a 4:2:0 frame -> their device form (host only, any
 * thread).  Chroma mask entries outside the 4:2:0 extent are dropped. */
int rav2d_gpu_pack_filter(const Rav2dAv2Filter *in, size_t n, Rav2dPackedFilter *out);
/* Segment maps of a segmented frame: FilterFrameParams' cur_segmap and
 * segmap_uv (decode.rs:285-310, 2355-2367), as deblock64_cols/rows read them
 * (deblock.rs:1623-2140).  segmap: one segment id (0..7) per luma 4x4 unit,
 * `stride` bytes per row, for the frame's ((W+7)>>3)*2 x ((H+7)>>3)*2 units;
 * segmap_uv: one id per chroma 4x4 unit (subsampled as the chroma planes),
 * uv_stride bytes per row.  Read by the deblocking thresholds when
 * params.seg_enabled is set: a segmented frame with deblocking on must have
 * its maps set before it is filtered (RAV2D_EINVAL otherwise).  Ids above 7
 * are rejected.  segmap NULL drops the maps.  Waits for the frame's earlier
 * launches like rav2d_gpu_frame_set_side_info. */
int rav2d_gpu_frame_set_segmap(Rav2dGpuFrame *f, const uint8_t *segmap, ptrdiff_t stride,
                               const uint8_t *segmap_uv, ptrdiff_t uv_stride, void *stream);
/* Replace the frame parameters (same size, bit depth and layout).  This and
 * rav2d_gpu_frame_set_side_info first wait for the frame's earlier
 * rav2d_gpu_filter_frame / apply_grain / run_pass launches and its grain
 * set-up, whose device tables they rewrite; a grain set-up built for the old
 * parameters is dropped.  A frame held by a batch run or a server job in
 * flight must not be changed (the batch and server sections say so). */
int rav2d_gpu_frame_set_params(Rav2dGpuFrame *f, const Rav2dFrameParams *params);
/* Device planes of the frame: which = 0 input F0, 1 output, 2 display. */
int rav2d_gpu_frame_planes(const Rav2dGpuFrame *f, int which, Rav2dPlanes *out);
/* Host <-> device copies of tightly packed planes (stride = width samples).
 * The upload writes the frame's visible samples only; the picture padding
 * (below) keeps what it holds: zero after rav2d_gpu_frame_create. */
int rav2d_gpu_frame_upload(Rav2dGpuFrame *f, const void *y, const void *u,
                           const void *v, void *stream);
/* Picture padding.  Like the reference's picture buffers (DefaultPicAllocator,
 * picture.rs:77-125: heights and widths aligned to 128), every plane of a
 * frame or picture of this library holds RAV2D_PICTURE_PAD rows below the
 * frame and, when the width is not a multiple of 128, RAV2D_PICTURE_PAD
 * samples right of it.  Deblocking reads them where a transform overhangs the
 * frame edge (filter_choice_bd looks up to 8 samples across an edge,
 * deblock.rs:81-186; the bottom-edge crop, :1572-1606, covers only luma
 * 32-long edges): a decoder's reconstruction of the overhanging blocks lives
 * there.  Planes handed to rav2d_gpu_filter_planes must carry the same
 * padding (rav2d_gpu_picture_alloc's do).  This upload copies the padding
 * too: src holds the frame plus pad_rows (<= 8) rows below and pad_cols
 * (<= 8) samples right of each plane, at src->stride bytes per row; columns
 * beyond the frame's own padding are skipped. */
#define RAV2D_PICTURE_PAD 8
int rav2d_gpu_frame_upload_padded(Rav2dGpuFrame *f, const Rav2dPlanes *src, int pad_rows, int pad_cols,
                                  void *stream);
int rav2d_gpu_frame_download(Rav2dGpuFrame *f, int which, void *y, void *u, void *v,
                             void *stream);

/* ---- frame-level entry points --------------------------------------------
 * rav2d_gpu_filter_frame replaces the PHASE-B loop (decode.rs:3597-3625):
 * input planes F0 -> output planes, all in-loop filters enabled by
 * params->inloop_filters.  rav2d_gpu_apply_grain replaces
 * apply_grain_to_picture_mt (decode.rs:4672): output planes -> display planes;
 * a no-op copy when the frame carries no grain (picture_has_grain). */
int rav2d_gpu_filter_frame(Rav2dGpuCtx *ctx, Rav2dGpuFrame *f, void *stream);
int rav2d_gpu_apply_grain(Rav2dGpuCtx *ctx, Rav2dGpuFrame *f, void *stream);

/* ---- zero-copy pictures ---------------------------------------------------
 * A PicAllocator backed by device memory (picture.rs:38-41, 77-125): pictures
 * in the library's plane layout (rav2d_gpu_frame_planes' strides for the same
 * size and depth: rows padded to 256 B, +256 B when the stride is a multiple
 * of 1 KiB, 8 spare rows), one allocation against the HBM budget.  I400
 * pictures have one plane (data[1], data[2] and stride[1] are 0).  With them
 * the decoder's reconstruction, the in-loop filters and film grain share
 * device pictures and the copies at decode.rs:4549 and :4681 go away. */
int rav2d_gpu_picture_alloc(Rav2dGpuCtx *ctx, int width, int height, int bitdepth, int layout, Rav2dPlanes *out);
int rav2d_gpu_picture_free(Rav2dGpuCtx *ctx, Rav2dPlanes *p);
/* The in-loop chain of rav2d_gpu_filter_frame on caller-owned device planes:
 * reads `in`, writes `out` (out == in filters in place), with the frame's
 * parameters, side info and intermediate planes.  Both plane sets need the
 * frame's strides and 8-byte-aligned plane pointers (RAV2D_EINVAL otherwise);
 * rav2d_gpu_picture_alloc pictures of the frame's size and depth qualify. */
int rav2d_gpu_filter_planes(Rav2dGpuCtx *ctx, Rav2dGpuFrame *f, const Rav2dPlanes *in, const Rav2dPlanes *out,
                            void *stream);
/* apply_grain_to_picture_mt (decode.rs:4672) on caller-owned device pictures:
 * dst = src with film grain `data` / `seed` (a copy when the data carries no
 * grain, picture_has_grain decode.rs:4641-4648).  Strides are multiples of 8
 * covering a row; any layout.  The grain templates, scaling LUTs and block offsets
 * live in a scratch area this call allocates and frees, stream-ordered
 * (hipMallocAsync / hipFreeAsync on `stream`). */
int rav2d_gpu_apply_grain_planes(Rav2dGpuCtx *ctx, const Rav2dPlanes *src, const Rav2dPlanes *dst, int width,
                                 int height, int bitdepth, int layout, const Rav2dFilmGrainData *data, uint32_t seed,
                                 void *stream);
/* No counterpart in the reference: rav2d_gpu_apply_grain_planes for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and output; the scratch
 * area is a static device array per GPU (572,748 bytes, held until the process
 * ends), so the call allocates nothing.  Pictures up to 8192 x 8192 luma
 * samples (RAV2D_EINVAL beyond).  Every call on one GPU shares that array:
 * calls that overlap on a GPU (other streams, other threads) race. */
int rav2d_gpu_apply_grain_planes_noalloc(Rav2dGpuCtx *ctx, const Rav2dPlanes *src, const Rav2dPlanes *dst,
                                         int width, int height, int bitdepth, int layout,
                                         const Rav2dFilmGrainData *data, uint32_t seed, void *stream);

/* ---- per-pass entry points (unit parity and per-kernel profiling) -------- */
#define RAV2D_PASS_DEBLOCK_V 0
#define RAV2D_PASS_DEBLOCK_H 1
#define RAV2D_PASS_CDEF 2
#define RAV2D_PASS_LR 3
#define RAV2D_PASS_GRAIN 4
int rav2d_gpu_run_pass(Rav2dGpuCtx *ctx, Rav2dGpuFrame *f, int pass, void *stream);

/* ---- timing with HIP events on the context stream ----------------------- */
/* When profiling is on, every pass launched through this context on its own
 * stream is bracketed by HIP events; rav2d_gpu_pass_times returns, per pass
 * (RAV2D_PASS_*), the summed milliseconds and the launch count since the last
 * reset (synchronizes). */
#define RAV2D_NUM_PASSES 5
int rav2d_gpu_set_profiling(Rav2dGpuCtx *ctx, int on);
int rav2d_gpu_pass_times(Rav2dGpuCtx *ctx, double *ms, int64_t *count, int reset);
int rav2d_gpu_timer_start(Rav2dGpuCtx *ctx);
/* Milliseconds between the last timer_start and now (synchronizes). */
int rav2d_gpu_timer_stop(Rav2dGpuCtx *ctx, float *ms);

/* ---- batched launches -------------------------------------------------------
 * No counterpart in the reference, which filters one frame at a time in its
 * PHASE-B loop (decode.rs:3597-3625) and then applies grain
 * (decode.rs:4672).  A batch runs exactly that per-frame work for many frames
 * of any size and bit depth with one launch per pass over all of them: the
 * output (and, with apply_grain, display) planes of every frame equal what
 * rav2d_gpu_filter_frame + rav2d_gpu_apply_grain write.  Frames are planned
 * when the batch is created and re-planned by the next run after their
 * parameters or side info change; a frame must stay alive while a batch
 * holding it exists, and must not be filtered by another call while a run
 * of the batch is in flight (they share the frame's intermediate planes). */
typedef struct Rav2dGpuBatch Rav2dGpuBatch;
typedef struct Rav2dBatchConfig {
    int32_t use_graph;                     /* capture once, replay as a hipGraph */
    int32_t apply_grain;                   /* also write the display planes */
    int32_t grid[RAV2D_NUM_PASSES];        /* workgroups of each pass launch
                                              (RAV2D_PASS_*; GRAIN = apply);
                                              0 = one per work item, else a
                                              persistent grid that loops */
    int32_t xcd_remap[RAV2D_NUM_PASSES];   /* give each XCD a contiguous range
                                              of work items */
    int32_t deblock_rows;                  /* 0: the two deblocking passes are
                                              separate launches; 32 or 64: one
                                              tile pass of 64 x deblock_rows
                                              tiles filters vertical, then
                                              horizontal edges (launch shape in
                                              grid / xcd_remap[DEBLOCK_V]) */
    int32_t mixed_depths;                  /* 1: one launch per pass over the
                                              frames of every bit depth (the
                                              bit depth dispatched per work
                                              item); 0: one launch per pass and
                                              bit depth, the 8-bit frames' chain
                                              on a parallel graph branch */
} Rav2dBatchConfig;
/* use_graph = 1, apply_grain = 1, one workgroup per item, no remap, separate
 * deblocking passes, one launch per bit depth. */
void rav2d_gpu_batch_default_config(Rav2dBatchConfig *cfg);
int rav2d_gpu_batch_create(Rav2dGpuCtx *ctx, Rav2dGpuFrame *const *frames, int n,
                           const Rav2dBatchConfig *cfg, Rav2dGpuBatch **out);
int rav2d_gpu_batch_destroy(Rav2dGpuBatch *b);
/* Enqueue one run of the whole batch on `stream` (NULL = context stream). */
int rav2d_gpu_batch_run(Rav2dGpuBatch *b, void *stream);
int rav2d_gpu_batch_get_config(const Rav2dGpuBatch *b, Rav2dBatchConfig *out);
int rav2d_gpu_batch_set_config(Rav2dGpuBatch *b, const Rav2dBatchConfig *cfg);
/* Times every launch shape of every pass (iters runs each, on the context
 * stream, which it synchronizes), keeps the fastest per pass and returns it. */
int rav2d_gpu_batch_autotune(Rav2dGpuBatch *b, int iters, Rav2dBatchConfig *chosen);
/* Work items, algorithmic bytes (one read + one write of the planes) and
 * frames of one pass's launch. */
int rav2d_gpu_batch_info(const Rav2dGpuBatch *b, int pass, int64_t *items, int64_t *bytes,
                         int *frames);
/* Luma pixels of all frames of the batch. */
int64_t rav2d_gpu_batch_pixels(const Rav2dGpuBatch *b);

/* ---- serving ----------------------------------------------------------------
 * No counterpart in the reference (a decoder library: one Decoder per stream,
 * frames filtered in decode order).  A server takes filter jobs from many
 * streams as they arrive.  A job names a device-resident frame whose input
 * planes, side info and parameters are set; it produces what
 * rav2d_gpu_filter_frame (+ rav2d_gpu_apply_grain with apply_grain) would.
 * A dispatcher thread owns max_inflight launch slots (one HIP stream each);
 * whenever a slot is free it gathers up to max_batch waiting jobs, oldest
 * first, into one batched launch per pass, so jobs that arrive while other
 * batches run are admitted into the next batch and overlap them on the
 * device (continuous batching).  A job waits while an earlier job of the same
 * frame is queued or running.  Memory admission is the context's HBM budget:
 * rav2d_gpu_frame_create returns RAV2D_EAGAIN once the budget cannot hold
 * another frame.  The caller must not change a frame while it has jobs. */
typedef struct Rav2dGpuServer Rav2dGpuServer;
typedef struct Rav2dServeConfig {
    int32_t max_batch;     /* jobs per batched launch (1 = one frame per launch) */
    int32_t max_inflight;  /* batches on the device at once, 1..8 */
    int32_t apply_grain;   /* also write the display planes */
    int32_t checksum;      /* hash each job's output and display planes */
} Rav2dServeConfig;
typedef struct Rav2dServeResult {
    uint64_t tag;
    double t_submit;       /* arrival, seconds on the server clock */
    double t_launch;       /* its batch was enqueued */
    double t_done;         /* its batch was seen complete */
    uint64_t checksum[2];  /* output, display planes (rav2d_gpu_frame_checksum) */
    int32_t status;        /* RAV2D_OK or the error of its batch */
    int32_t batch_frames;  /* jobs in its batch */
} Rav2dServeResult;
/* max_batch 16, max_inflight 2, apply_grain 1, checksum 0.  Each slot owns
 * one HIP stream; the slots share the context's auxiliary stream for grain
 * set-ups, and the argument buffers of a max_batch batch are allocated per
 * slot by rav2d_gpu_server_create (the dispatcher allocates nothing). */
void rav2d_gpu_server_default_config(Rav2dServeConfig *cfg);
int rav2d_gpu_server_create(Rav2dGpuCtx *ctx, const Rav2dServeConfig *cfg, Rav2dGpuServer **out);
/* Finishes every submitted job, then stops the dispatcher. */
int rav2d_gpu_server_destroy(Rav2dGpuServer *s);
/* Seconds since the server was created. */
double rav2d_gpu_server_clock(const Rav2dGpuServer *s);
/* Queue a job; t_arrival < 0 stamps it with the current clock. */
int rav2d_gpu_server_submit(Rav2dGpuServer *s, Rav2dGpuFrame *f, uint64_t tag, double t_arrival);
/* Up to max completed jobs (waiting up to wait_ms for the first); returns the
 * number written, or a negative error code. */
int rav2d_gpu_server_poll(Rav2dGpuServer *s, Rav2dServeResult *out, int max, int wait_ms);
/* Blocks until every submitted job has completed. */
int rav2d_gpu_server_drain(Rav2dGpuServer *s);
/* paused != 0: the dispatcher launches no new batch (batches already on the
 * device finish; queued jobs wait) until the server is resumed with 0. */
int rav2d_gpu_server_pause(Rav2dGpuServer *s, int paused);
/* Cancel a job still queued (not yet in a batch): its result comes back from
 * rav2d_gpu_server_poll with status RAV2D_ECANCELED.  RAV2D_ENOENT when no
 * queued job has the tag (already launched, done, or never submitted). */
int rav2d_gpu_server_cancel(Rav2dGpuServer *s, uint64_t tag);
/* Simulated streams: stream i owns frames[i*slots .. i*slots + slots-1],
 * set up by the caller.  Each stream's frames arrive as a Poisson process of
 * rate fps (seeded); job j filters its stream's next slot in turn.  Submits
 * n_jobs jobs at their arrival times and returns when all completed;
 * results[j] is job j's record, job_stream/job_slot (optional) its stream
 * and slot. */
int rav2d_gpu_server_simulate(Rav2dGpuServer *s, Rav2dGpuFrame *const *frames, int n_streams,
                              int slots, double fps, int n_jobs, uint64_t seed,
                              Rav2dServeResult *results, int32_t *job_stream, int32_t *job_slot);
/* Order-independent 64-bit hash of a frame's planes (which = 0 input,
 * 1 output, 2 display): the sum over samples of splitmix64(plane << 60 |
 * y << 40 | x << 20 | value).  Synchronizes the context stream. */
int rav2d_gpu_frame_checksum(Rav2dGpuFrame *f, int which, uint64_t *out);

/* Achievable HBM bandwidth: a streaming copy of `bytes` (multiple of 16,
 * >= 4096) between two fresh buffers, iters times on the context stream;
 * *gbps = (read + write bytes) / time.  No counterpart in the reference: a
 * measurement utility. */
int rav2d_gpu_bandwidth_probe(Rav2dGpuCtx *ctx, uint64_t bytes, int iters, double *gbps);

/* ---- per-family DSP entries -------------------------------------------------
 * The function-table entries of rav2d's DSP contexts for the filter stage
 * (DeblockDSPContext, CdefDSPContext, CcsoDSPContext,
 * LoopRestorationDSPContext, FilmGrainDSPContext; crates/rav2d/src/dsp.rs:
 * 9-14, 86-109), one GPU launch over one call's unit each, for table-level
 * drop-in and unit parity; the frame-level entries above are the fast path.
 * Arguments follow the reference's kernels one for one.  Pixel buffers
 * (dst/src, the left/top/bottom line buffers, luma rows, grain templates,
 * scaling LUTs, index and error maps) are DEVICE memory: samples are uint8_t
 * at 8-bit and uint16_t at 10-bit and strides are in bytes.  Small side arrays
 * (edge masks, thresholds, lossless masks, filters, class LUTs, film-grain
 * data) are HOST memory, copied into the launch.  A call is enqueued on
 * `stream` (NULL = the context stream) and returns at once.  In-place calls
 * read their whole input before writing (as the reference's line backups
 * do).  RAV2D_EINVAL for a bit depth other than 8/10 or a unit larger than
 * the reference's (stated per entry). */

/* CDEF_HAVE_* / LR_HAVE_* edge bits (cdef.rs, looprestoration.rs) */
#define RAV2D_EDGE_LEFT 1
#define RAV2D_EDGE_RIGHT 2
#define RAV2D_EDGE_TOP 4
#define RAV2D_EDGE_BOTTOM 8

/* DeblockDSPContext.deblock_sb[dir][uv]: deblock_{h,v}_sb64{y,uv}_bd
 * (deblock.rs:303-551).  The 4-line segments of one edge line: segment i
 * (bit i of the level masks) starts 4 i samples along the edge from dst, which
 * is the first sample on the edge's positive side.  dir 0 = vertical edges
 * (the "h" walkers: segments go down the rows), 1 = horizontal edges (the "v"
 * walkers).  vmask: the level masks (4 luma, 3 chroma); ll_mask[0] / [1]:
 * lossless negative / positive side; q_thr, side_thr: one per segment (16);
 * edge: clamp the negative side's width (6 luma, 2 chroma). */
int rav2d_gpu_deblock_sb(Rav2dGpuCtx *ctx, int dir, int uv, void *dst, ptrdiff_t stride, const uint16_t *vmask,
                         const uint16_t ll_mask[2], const uint16_t q_thr[16], const uint16_t side_thr[16], int edge,
                         int bitdepth, void *stream);

/* CdefDSPContext.dir: cdef_find_dir_bd (cdef.rs:134-216) of the 8x8 block at
 * img; the direction (0..7) and variance go to dir_out / var_out (device). */
int rav2d_gpu_cdef_dir(Rav2dGpuCtx *ctx, const void *img, ptrdiff_t stride, int bitdepth, int32_t *dir_out,
                       uint32_t *var_out, void *stream);

/* CdefDSPContext.fb[]: cdef_filter_block (cdef.rs:320-460), the w x h block
 * at dst (w, h in {4, 8}) filtered in place.  left: h rows of 2 samples (the
 * columns left of the block); top / bottom: 2 rows each at dst's stride,
 * pointing at the block's column 0 (columns -2 .. w + 1 readable); edges:
 * RAV2D_EDGE_* of the neighbours that exist. */
int rav2d_gpu_cdef_filter(Rav2dGpuCtx *ctx, void *dst, ptrdiff_t stride, const void *left, const void *top,
                          const void *bottom, int pri_strength, int sec_strength, int dir, int damping, int w, int h,
                          int edges, int bitdepth, void *stream);

/* CcsoDSPContext.prep[plane]: ccso_prep (ccso.rs:26-186), the band / edge
 * class byte of each of the w x h samples of a plane block (its luma, w << ss_hor
 * by h << ss_ver, at most 64 x 64: one superblock column, as the reference's)
 * from its collocated luma: src (src_stride), left (rows of 2 luma samples),
 * top / bottom (2 luma rows each at src_stride, at the block's column 0);
 * ext_filter indexes CCSO_POS (0..6); dst: bytes at dst_stride. */
int rav2d_gpu_ccso_prep(Rav2dGpuCtx *ctx, uint8_t *dst, ptrdiff_t dst_stride, const void *src, ptrdiff_t src_stride,
                        const void *left, const void *top, const void *bottom, int max_band_log2, int ext_filter,
                        int quant_step, uint32_t edge_clf, int bo_only, int edges, int w, int h, int ss_hor,
                        int ss_ver, int bitdepth, void *stream);

/* CcsoDSPContext.add: ccso_add (ccso.rs:215-254), dst += offset_lut[the
 * nibble of offset_idxs the class byte selects], clipped, except lossless 4x4
 * blocks (ll_mask: h / 4 rows, word 0 = bit per 4x4 column; host).
 * offset_idxs: 64 bytes, offset_lut: 8 (host); w <= 256, h <= 64. */
int rav2d_gpu_ccso_add(Rav2dGpuCtx *ctx, void *dst, ptrdiff_t dst_stride, const uint8_t *idx, ptrdiff_t idx_stride,
                       const uint8_t offset_idxs[64], const int8_t offset_lut[8], int w, int h,
                       const uint16_t (*ll_mask)[4], int bitdepth, void *stream);

/* WienerParamsSingle / WienerParamsMulti / WienerParamsC
 * (looprestoration.rs:1716-1740).  Single: filter = 16 (luma) or 18 (chroma)
 * host taps; luma / luma_top / luma_bottom (device) and stride, ss_ver,
 * ss_hor, ds_flt are the chroma kernel's collocated luma.  Multi: filters =
 * host [16][18] int8 class filters (NS multi) or [64][13] int16 pretrained
 * filters (PC-Wiener); subclass_lut: host 256 bytes; noskip_mask: host, one
 * 16-bit row per 4 sample rows of the stripe plus one. */
typedef struct Rav2dWienerParamsSingle {
    const int8_t *filter;
    const void *luma;
    const void *luma_top;
    const void *luma_bottom;
    ptrdiff_t stride;
    int32_t ss_ver;
    int32_t ss_hor;
    int32_t ds_flt;
} Rav2dWienerParamsSingle;
typedef struct Rav2dWienerParamsMulti {
    const uint8_t *filters;
    const uint8_t *subclass_lut;
    const uint16_t *noskip_mask;
    int32_t base_q;
} Rav2dWienerParamsMulti;
typedef union Rav2dWienerParams {
    Rav2dWienerParamsSingle single;
    Rav2dWienerParamsMulti multi;
} Rav2dWienerParams;

/* LoopRestorationDSPContext entries through WienerFilterFn
 * (looprestoration.rs:1742-1753): kind 0 ns_wiener_single (luma, :524-713),
 * 1 ns_wiener_single (chroma with its cross-component luma taps, 4:2:0,
 * :1145-1430), 2 ns_wiener_multi (:715-1100), 3 pc_wiener (:1102-1141).  The
 * stripe of one restoration unit at dst (w <= 64, h <= 64 samples, multiples
 * of 4); left: h rows of 6 samples; top / bottom: the 2 rows above / below at
 * dst's stride; ll_mask: h / 4 + 1 host rows of 4 words (word 0 = lossless bit
 * per 4x4 column). */
int rav2d_gpu_lr_wiener(Rav2dGpuCtx *ctx, int kind, void *dst, ptrdiff_t dst_stride, const void *left,
                        const void *top, const void *bottom, int w, int h, const Rav2dWienerParams *params,
                        uint8_t edges, const uint16_t (*ll_mask)[4], int bitdepth, void *stream);
/* GdfPrepFn (looprestoration.rs:1755-1768): gdf_prep (:1432-1643), the int8
 * error of each sample of the w x h stripe at p (w <= 64, h <= 64, even); left:
 * h rows of 6 samples, top / bottom: the 2 rows above / below at p's stride. */
int rav2d_gpu_gdf_prep(Rav2dGpuCtx *ctx, int8_t *dst, ptrdiff_t dst_stride, const void *p, ptrdiff_t stride,
                       const void *left, const void *top, const void *bottom, int w, int h, int ref_dst_idx,
                       int qp_idx, uint8_t edges, int bitdepth, void *stream);
/* GdfAddFn (looprestoration.rs:1770-1779): gdf_add (:273-329), p += the
 * scaled error, clipped, except lossless 4x4 blocks (w, h <= 64). */
int rav2d_gpu_gdf_add(Rav2dGpuCtx *ctx, void *p, ptrdiff_t stride, const int8_t *err, ptrdiff_t err_stride, int w,
                      int h, int scale, const uint16_t (*ll_mask)[4], int bitdepth, void *stream);

/* FilmGrainDSPContext.generate_grain_y / generate_grain_uv[layout]
 * (filmgrain.rs:115-229): the 73 x 82 luma template, or a chroma template
 * built from the luma template buf_y: 38 x 44 for (subx, suby) = (1, 1)
 * (4:2:0), 73 x 44 for (1, 0) (4:2:2), 73 x 82 for (0, 0) (4:4:4); rows of
 * 82 entries; buffers int16_t, device. */
int rav2d_gpu_generate_grain_y(Rav2dGpuCtx *ctx, int16_t *buf, const Rav2dFilmGrainData *data, uint32_t seed,
                               int bitdepth, void *stream);
int rav2d_gpu_generate_grain_uv(Rav2dGpuCtx *ctx, int16_t *buf, const int16_t *buf_y,
                                const Rav2dFilmGrainData *data, uint32_t seed, int uv, int subx, int suby,
                                int bitdepth, void *stream);
/* FilmGrainDSPContext.fgy_32x32xn / fguv_32x32xn[layout] (filmgrain.rs:248-592):
 * one row band (bh rows, band row_num) of a plane, pw samples wide; dst, src
 * at stride; scaling: the plane's 1 << bitdepth LUT (device); grain_lut: the
 * plane's template (device, rows of 82); fguv also reads the band's luma rows
 * (luma_row at luma_stride), is_id = mc_identity, (sx, sy) = (1, 1) 4:2:0,
 * (1, 0) 4:2:2 or (0, 0) 4:4:4. */
int rav2d_gpu_fgy_32x32xn(Rav2dGpuCtx *ctx, void *dst, const void *src, ptrdiff_t stride,
                          const Rav2dFilmGrainData *data, uint32_t seed, int pw, const uint8_t *scaling,
                          const int16_t *grain_lut, int bh, int row_num, int bitdepth, void *stream);
int rav2d_gpu_fguv_32x32xn(Rav2dGpuCtx *ctx, void *dst, const void *src, ptrdiff_t stride,
                           const Rav2dFilmGrainData *data, uint32_t seed, int pw, const uint8_t *scaling,
                           const int16_t *grain_lut, int bh, int row_num, const void *luma_row,
                           ptrdiff_t luma_stride, int uv, int is_id, int sx, int sy, int bitdepth, void *stream);

/* sizeof() of the ABI structs, so bindings can check their mirrors:
 * 0 Rav2dAv2Filter, 1 Rav2dAv2Restoration, 2 Rav2dFrameParams,
 * 3 Rav2dFilmGrainData, 4 Rav2dPlanes, 5 Rav2dBatchConfig, 6 Rav2dServeConfig,
 * 7 Rav2dServeResult, 8 Rav2dWienerParams, 9 Rav2dPackedFilter; 0 for an
 * unknown index. */
size_t rav2d_gpu_struct_size(int which);

const char *rav2d_gpu_version(void);
/* Text of the last error recorded on this thread. */
const char *rav2d_gpu_last_error(void);

/* Fault injection for tests of the error paths (no counterpart in the
 * reference): the next `count` device allocations of the library (frames,
 * rav2d_gpu_alloc, batch argument blocks; what = RAV2D_FAIL_ALLOC) or the
 * next `count` filter launches (the per-frame chain and batch / server
 * enqueues; what = RAV2D_FAIL_LAUNCH) fail on the host as if the HIP call had
 * failed, before anything reaches the device.  count 0 clears it. */
#define RAV2D_FAIL_ALLOC 0
#define RAV2D_FAIL_LAUNCH 1
int rav2d_gpu_debug_inject_failure(int what, int count);

#ifdef __cplusplus
}
#endif
#endif /* RAV2D_GPU_H */
