/*
 * Oracle frame driver: the per-superblock-row order of filter_sbrow
 * (crates/rav2d/src/decode.rs:3718-4058) and the output grain of
 * apply_grain_to_picture_mt (decode.rs:4641-4775).  TEST INFRASTRUCTURE ONLY.
 */
#include "oracle_common.h"
#include "oracle.h"

namespace oracle {
template <typename P>
void deblock_sbrow_impl(const Rav2dFrameParams *p, const Geom &g, Rav2dAv2Filter *mask, size_t n_mask,
                        Frame<P> &f, int sby, bool cols_only);
}

using namespace oracle;

template <typename P>
static int run_frame(const Rav2dFrameParams *p, const Rav2dAv2Filter *mask_in, size_t n_mask,
                     const Rav2dAv2Restoration *lr, size_t n_lr, P *y, P *u, P *v, ptrdiff_t ys, ptrdiff_t uvs,
                     uint32_t stop_mask) {
    (void)lr;
    (void)n_lr;
    Geom g(p);
    /* the bottom-edge crop mutates the mask: work on a copy (deblock.rs:1569-1606) */
    std::vector<Rav2dAv2Filter> mask(mask_in, mask_in + n_mask);
    Frame<P> f{y, u, v, ys, uvs};
    uint32_t inloop = p->inloop_filters & stop_mask;
    bool deblock_on = (inloop & RAV2D_INLOOPFILTER_DEBLOCK) && (p->db_level_y[0] || p->db_level_y[1]);
    for (int sby = 0; sby < g.sbh; sby++) {
        if (deblock_on)
            deblock_sbrow_impl(p, g, mask.data(), mask.size(), f, sby, (stop_mask & ORACLE_DEBLOCK_COLS_ONLY) != 0);
    }
    return 0;
}

extern "C" int oracle_filter_frame(const Rav2dFrameParams *p, const Rav2dAv2Filter *mask, size_t n_mask,
                                   const Rav2dAv2Restoration *lr, size_t n_lr, void *y, void *u, void *v,
                                   ptrdiff_t y_stride, ptrdiff_t uv_stride, uint32_t inloop_mask) {
    if (!p || !mask || p->layout != RAV2D_PIXEL_LAYOUT_I420 || p->seg_enabled) return RAV2D_EINVAL;
    if (p->bitdepth == 8)
        return run_frame<uint8_t>(p, mask, n_mask, lr, n_lr, (uint8_t *)y, (uint8_t *)u, (uint8_t *)v, y_stride,
                                  uv_stride, inloop_mask);
    if (p->bitdepth == 10)
        return run_frame<uint16_t>(p, mask, n_mask, lr, n_lr, (uint16_t *)y, (uint16_t *)u, (uint16_t *)v,
                                   y_stride / 2, uv_stride / 2, inloop_mask);
    return RAV2D_EINVAL;
}
