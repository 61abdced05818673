/* Oracle loop-restoration declarations (TEST INFRASTRUCTURE ONLY). */
#pragma once
#include "oracle_common.h"

namespace oracle {

static const int LR_RESTORE_Y = 1, LR_RESTORE_U = 2, LR_RESTORE_V = 4;

/* LrContext, looprestoration.rs:1785-1818, for one frame */
template <typename P> struct LrCtx {
    const Rav2dFrameParams *p;
    const Geom *g;
    const Rav2dAv2Filter *mask;
    size_t n_mask;
    const Rav2dAv2Restoration *lr;
    size_t n_lr;
    std::vector<P> lr_db[3];       /* deblocked-line store, 20 rows per plane */
    const P *luma_snapshot;        /* pre-luma-LR luma of the current sbrow */
    ptrdiff_t ys, uvs;             /* strides in samples */
    uint32_t inloop;
    int base_q, restore_planes;
    const uint8_t *ns_subclass_lut, *pc_subclass_lut;
    const int16_t (*pc_filters)[13];
};

int get_qval_given_tskip(int qstep, int tskip, int i, int bdm8);
template <typename P> void lr_init(LrCtx<P> &c, const Rav2dFrameParams *p);
template <typename P> void copy_db(LrCtx<P> &c, const Frame<P> &f, int sby);
template <typename P> void lr_sbrow_apply(const LrCtx<P> &c, Frame<P> &f, int sby);

}  // namespace oracle
