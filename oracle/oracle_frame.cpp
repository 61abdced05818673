/*
 * Oracle frame driver: the per-superblock-row order of filter_sbrow
 * (crates/rav2d/src/decode.rs:3718-4058) and the output grain of
 * apply_grain_to_picture_mt (decode.rs:4641-4775).  TEST INFRASTRUCTURE ONLY.
 */
#include "oracle_common.h"
#include "oracle.h"
#include "oracle_cdef.h"
#include "oracle_lr.h"
#include "../rav2d_amd/csrc/tables.h"

namespace oracle {
template <typename P>
void deblock_sbrow_impl(const Rav2dFrameParams *p, const Geom &g, Rav2dAv2Filter *mask, size_t n_mask,
                        Frame<P> &f, int sby, bool cols_only);
}

using namespace oracle;

template <typename P>
static int run_frame(const Rav2dFrameParams *p, const Rav2dAv2Filter *mask_in, size_t n_mask,
                     const Rav2dAv2Restoration *lr, size_t n_lr, P *y, P *u, P *v, ptrdiff_t ys, ptrdiff_t uvs,
                     uint32_t stop_mask, const uint8_t *seg, ptrdiff_t seg_stride, const uint8_t *seg_uv,
                     ptrdiff_t uv_seg_stride) {
    (void)lr;
    (void)n_lr;
    Geom g(p);
    if (p->seg_enabled) {
        g.seg = seg;
        g.seg_uv = seg_uv;
        g.seg_stride = seg_stride;
        g.uv_seg_stride = uv_seg_stride;
    }
    /* the bottom-edge crop mutates the mask: work on a copy (deblock.rs:1569-1606) */
    std::vector<Rav2dAv2Filter> mask(mask_in, mask_in + n_mask);
    Frame<P> f{y, u, v, ys, uvs};
    uint32_t inloop = p->inloop_filters & stop_mask;
    bool deblock_on = (inloop & RAV2D_INLOOPFILTER_DEBLOCK) && (p->db_level_y[0] || p->db_level_y[1]);
    /* CDEF (+CCSO) set-up, decode.rs:3858-3913 */
    bool cdef_stage = p->seq_cdef && (inloop & (RAV2D_INLOOPFILTER_CDEF | RAV2D_INLOOPFILTER_CCSO));
    CdefBrowParams cp;
    memset(&cp, 0, sizeof(cp));
    cp.bw = g.bw;
    cp.bh = g.bh;
    cp.damping = p->cdef_damping;
    cp.on_skip_tx = p->cdef_on_skiptx != 0;
    cp.cdef_on = (inloop & RAV2D_INLOOPFILTER_CDEF) != 0;
    cp.have_chroma = p->layout != RAV2D_PIXEL_LAYOUT_I400;
    cp.layout = p->layout;
    /* decode.rs:3868-3869 */
    cp.any_lossless = false;
    for (int i = 0; i < 8; i++) cp.any_lossless = cp.any_lossless || (p->seg_enabled && p->seg_lossless[i]);
    cp.ccso_on = (inloop & RAV2D_INLOOPFILTER_CCSO) && (p->ccso[0].enabled || p->ccso[1].enabled || p->ccso[2].enabled);
    cp.y_strength = p->cdef_y_strength;
    cp.uv_strength = p->cdef_uv_strength;
    cp.mask = mask.data();
    cp.n_mask = mask.size();
    cp.sb256w = g.sb256w;
    for (int pl = 0; pl < 3; pl++) {
        const Rav2dCcsoPlane &c = p->ccso[pl];
        cp.ccso[pl].max_band_log2 = c.max_band_log2;
        cp.ccso[pl].ext_filter = c.ext_filter_support;
        cp.ccso[pl].quant_step = R2T_CCSO_QUANT_SZ[c.scale_idx][c.quant_idx];
        cp.ccso[pl].edge_clf = c.edge_clf;
        cp.ccso[pl].bo_only = c.bo_only != 0;
        cp.ccso[pl].scale_idx = c.scale_idx;
        memcpy(cp.ccso[pl].filter_off, c.filter_off, 64);
    }
    CdefLines<P> lines;
    for (int b = 0; b < 2; b++) {
        lines.bank[b][0].assign((size_t)(2 * ys), 0);
        lines.bank[b][1].assign((size_t)(2 * uvs), 0);
        lines.bank[b][2].assign((size_t)(2 * uvs), 0);
    }
    lines.toggle = 0;
    /* loop restoration set-up, decode.rs:3812-3853, 3999-4057 */
    LrCtx<P> lc;
    lc.p = p;
    lc.g = &g;
    lc.mask = mask.data();
    lc.n_mask = mask.size();
    lc.lr = lr;
    lc.n_lr = lr ? n_lr : 0;
    lc.ys = ys;
    lc.uvs = uvs;
    lc.inloop = inloop;
    lr_init(lc, p);
    for (int pl = 0; pl < 3; pl++) lc.lr_db[pl].assign((size_t)(20 * (pl ? uvs : ys)), 0);
    bool lr_stage = lc.restore_planes && (inloop & (RAV2D_INLOOPFILTER_WIENER | RAV2D_INLOOPFILTER_GDF));
    std::vector<P> snapshot;
    for (int sby = 0; sby < g.sbh; sby++) {
        if (deblock_on)
            deblock_sbrow_impl(p, g, mask.data(), mask.size(), f, sby, (stop_mask & ORACLE_DEBLOCK_COLS_ONLY) != 0);
        /* copy_db (decode.rs:3818-3853) */
        if (lr_stage) copy_db(lc, f, sby);
        if (cdef_stage) {
            int start = sby * g.sb_step;
            if (sby > 0) {
                /* cross-row seam with the previous mask row (decode.rs:3914-3956) */
                CdefBrowParams prev = cp;
                prev.row = (size_t)(((sby - 1) >> (2 - g.sb128)) * g.sb256w);
                cdef_brow(prev, f, lines, start - 2, start, g.bitdepth);
            }
            CdefBrowParams cur = cp;
            cur.row = (size_t)((sby >> (2 - g.sb128)) * g.sb256w);
            int n_blks = g.sb_step - 2 * (sby + 1 < g.sbh);
            int end = imin(start + n_blks, g.bh);
            cdef_brow(cur, f, lines, start, end, g.bitdepth);
        }
        if (lr_stage) {
            /* chroma reads the pre-luma-LR luma of this sbrow (decode.rs:4004-4007) */
            snapshot.assign(f.y, f.y + (size_t)ys * ((g.bh * 4) + 0));
            lc.luma_snapshot = snapshot.data();
            lr_sbrow_apply(lc, f, sby);
        }
    }
    return 0;
}

extern "C" int oracle_filter_frame_seg(const Rav2dFrameParams *p, const Rav2dAv2Filter *mask, size_t n_mask,
                                       const Rav2dAv2Restoration *lr, size_t n_lr, void *y, void *u, void *v,
                                       ptrdiff_t y_stride, ptrdiff_t uv_stride, uint32_t inloop_mask,
                                       const uint8_t *seg, ptrdiff_t seg_stride, const uint8_t *seg_uv,
                                       ptrdiff_t uv_seg_stride) {
    if (!p || !mask || p->layout < RAV2D_PIXEL_LAYOUT_I400 || p->layout > RAV2D_PIXEL_LAYOUT_I444) return RAV2D_EINVAL;
    if (p->seg_enabled && (!seg || !seg_uv)) return RAV2D_EINVAL;
    if (p->bitdepth == 8)
        return run_frame<uint8_t>(p, mask, n_mask, lr, n_lr, (uint8_t *)y, (uint8_t *)u, (uint8_t *)v, y_stride,
                                  uv_stride, inloop_mask, seg, seg_stride, seg_uv, uv_seg_stride);
    if (p->bitdepth == 10 || p->bitdepth == 12)
        return run_frame<uint16_t>(p, mask, n_mask, lr, n_lr, (uint16_t *)y, (uint16_t *)u, (uint16_t *)v,
                                   y_stride / 2, uv_stride / 2, inloop_mask, seg, seg_stride, seg_uv, uv_seg_stride);
    return RAV2D_EINVAL;
}

extern "C" int oracle_filter_frame(const Rav2dFrameParams *p, const Rav2dAv2Filter *mask, size_t n_mask,
                                   const Rav2dAv2Restoration *lr, size_t n_lr, void *y, void *u, void *v,
                                   ptrdiff_t y_stride, ptrdiff_t uv_stride, uint32_t inloop_mask) {
    return oracle_filter_frame_seg(p, mask, n_mask, lr, n_lr, y, u, v, y_stride, uv_stride, inloop_mask, nullptr, 0,
                                   nullptr, 0);
}
