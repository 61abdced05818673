/* Oracle CDEF/CCSO declarations (TEST INFRASTRUCTURE ONLY). */
#pragma once
#include "oracle_common.h"

namespace oracle {

/* CcsoPlaneCfg, cdef.rs:562-571, built as build_ccso_plane_cfg (decode.rs:4117-4133) */
struct CcsoCfg {
    uint32_t max_band_log2;
    int ext_filter;
    int quant_step;
    uint32_t edge_clf;
    bool bo_only;
    int scale_idx;
    uint8_t filter_off[64];
};

/* CdefBrowParams, cdef.rs:581-603: the mask row is read through sb(). */
struct CdefBrowParams {
    int bw, bh, damping;
    bool on_skip_tx, cdef_on, ccso_on;
    bool any_lossless; /* a segment is lossless: the masks' lossless 4x4 units are skipped */
    bool have_chroma; /* false for I400 (cdef.rs:728): chroma is neither backed up nor filtered */
    int layout;       /* RAV2D_PIXEL_LAYOUT_*: chroma subsampling and the 4:2:2 direction remap (cdef.rs:722-725) */
    const uint8_t *y_strength, *uv_strength;
    const Rav2dAv2Filter *mask;
    size_t n_mask, row;
    int sb256w;
    CcsoCfg ccso[3];
    const Rav2dAv2Filter *sb(int sb256x) const {
        if (sb256x >= sb256w || row + sb256x >= n_mask) return nullptr;
        return &mask[row + sb256x];
    }
};

/* cdef_line[2][3] banks (2 rows each at the plane stride) + toggle,
 * LoopFilterState internal.rs:151-187 */
template <typename P> struct CdefLines {
    std::vector<P> bank[2][3];
    int toggle = 0;
};

template <typename P>
void cdef_brow(const CdefBrowParams &p, Frame<P> &f, CdefLines<P> &L, int by_start, int by_end, int bitdepth);

int cdef_adjust_strength(int strength, uint32_t var);
template <typename P> int cdef_find_dir(const P *img, ptrdiff_t stride, uint32_t *var, int bdm8);

}  // namespace oracle
