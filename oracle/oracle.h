/*
 * C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY: loaded by the
 * tests, the smoke check and the benchmark's CPU-baseline leg).
 */
#pragma once
#include "../include/rav2d_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Extra bit of `inloop_mask`: stop deblocking after the vertical-edge pass. */
#define ORACLE_DEBLOCK_COLS_ONLY (1u << 8)

/* Run the in-loop filter chain in place on host planes (strides in bytes),
 * in rav2d's per-superblock-row order.  `inloop_mask` is ANDed with
 * params->inloop_filters (use RAV2D_INLOOPFILTER_ALL for the full chain). */
int oracle_filter_frame(const Rav2dFrameParams *p, const Rav2dAv2Filter *mask, size_t n_mask,
                        const Rav2dAv2Restoration *lr, size_t n_lr, void *y, void *u, void *v,
                        ptrdiff_t y_stride, ptrdiff_t uv_stride, uint32_t inloop_mask);
/* The same for a segmented frame: seg / seg_uv are the segment maps of the
 * luma / chroma 4x4 units (strides in bytes), read when p->seg_enabled. */
int oracle_filter_frame_seg(const Rav2dFrameParams *p, const Rav2dAv2Filter *mask, size_t n_mask,
                            const Rav2dAv2Restoration *lr, size_t n_lr, void *y, void *u, void *v,
                            ptrdiff_t y_stride, ptrdiff_t uv_stride, uint32_t inloop_mask, const uint8_t *seg,
                            ptrdiff_t seg_stride, const uint8_t *seg_uv, ptrdiff_t uv_seg_stride);

/* Film grain on the output copy: dst = src with grain (planes of the display
 * copy), as apply_grain_to_picture_mt; a plain copy when the frame has none. */
int oracle_apply_grain(const Rav2dFrameParams *p, const void *src_y, const void *src_u, const void *src_v,
                       void *dst_y, void *dst_u, void *dst_v, ptrdiff_t y_stride, ptrdiff_t uv_stride);
int oracle_picture_has_grain(const Rav2dFrameParams *p);

/* Known-answer hooks into the restated helpers. */
int oracle_dq_lookup(int qidx);
uint32_t oracle_deblock_quant_thr(int hbd, int qidx);
uint32_t oracle_deblock_side_thr(int hbd, int qidx);

#ifdef __cplusplus
}
#endif
