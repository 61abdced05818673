/*
 * CPU oracle for the rav2d post-reconstruction filter stage.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a scalar restatement of rav2d's filter
 * chain, written to check the GPU path and to time a CPU baseline.  Only the
 * tests, the smoke check and the benchmark's CPU-baseline leg load it; the
 * product library never links or calls it.
 *
 * It follows the reference's own order of work: per superblock row, deblock
 * columns, deblock rows, line backups, CDEF with CCSO (including the
 * cross-row seam), then loop restoration, exactly as filter_sbrow
 * (crates/rav2d/src/decode.rs:3718-4058) does, so that comparing it with the
 * GPU's whole-frame passes also checks that the two orders agree.
 */
#pragma once
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <vector>
#include <algorithm>

#include "../include/rav2d_gpu.h"

namespace oracle {

/* intops.rs:1-80 */
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int iclip(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int apply_sign(int v, int s) { return s < 0 ? -v : v; }
static inline int ulog2(uint32_t v) { return 31 ^ __builtin_clz(v); }

/* quantizer.rs:31-47 */
int dq_lookup(int qidx);

/* Frame geometry of decode.rs:4244-4251. */
struct Geom {
    int bw, bh, sb256w, sb256h, sb128, sb_shift, sb_step, sbh;
    int ss_hor, ss_ver, hbd, bitdepth, bdmax, layout;
    /* segment maps (cur_segmap / segmap_uv of decode.rs:285-310, 2355-2367);
     * null on frames without segmentation (the PLACEHOLDER_SEGMAP path) */
    const uint8_t *seg = nullptr, *seg_uv = nullptr;
    ptrdiff_t seg_stride = 0, uv_seg_stride = 0;
    Geom(const Rav2dFrameParams *p) {
        bw = ((p->width + 7) >> 3) << 1;
        bh = ((p->height + 7) >> 3) << 1;
        sb256w = (bw + 63) >> 6;
        sb256h = (bh + 63) >> 6;
        sb128 = p->sb128 ? 1 : 0;
        sb_shift = 4 + sb128;
        sb_step = 16 << sb128;
        sbh = (bh + sb_step - 1) >> sb_shift;
        layout = p->layout;
        ss_ver = p->layout == RAV2D_PIXEL_LAYOUT_I420;
        ss_hor = p->layout == RAV2D_PIXEL_LAYOUT_I420 || p->layout == RAV2D_PIXEL_LAYOUT_I422;
        bitdepth = p->bitdepth;
        hbd = (p->bitdepth - 8) >> 1; /* SequenceHeader.hbd: 0, 1, 2 for 8, 10, 12 bit */
        bdmax = (1 << p->bitdepth) - 1;
    }
};

/* Planes addressed in samples. */
template <typename P> struct Frame {
    P *y, *u, *v;
    ptrdiff_t ys, uvs; /* strides in samples */
};

/* deblock (oracle_deblock.cpp) */
template <typename P>
void deblock_sbrow(const Rav2dFrameParams *p, const Geom &g, Rav2dAv2Filter *mask,
                   Frame<P> &f, int sby);

uint32_t deblock_quant_thr(int hbd, int qidx);
uint32_t deblock_side_thr(int hbd, int qidx);

/* CDEF + CCSO (oracle_cdef.cpp) */
template <typename P> struct CdefState;

/* loop restoration (oracle_lr.cpp) */

} // namespace oracle
