/*
 * Oracle: AV2 deblocking, restated from crates/rav2d/src/deblock.rs and
 * crates/rav2d/src/lf_mask.rs.  TEST INFRASTRUCTURE ONLY (see oracle_common.h).
 * On frames without segmentation the segment id of every 4x4 block is 0 (the
 * reference's PLACEHOLDER_SEGMAP path, deblock.rs:1463, 1675-1679).
 */
#include "oracle_common.h"
#include "../rav2d_amd/csrc/tables.h"

namespace oracle {

/* deblock.rs:6-11 */
static const int8_t MAX_WIDTH_Y[4] = {1, 3, 6, 8};
static const int8_t MAX_WIDTH_UV[3] = {1, 3, 4};
static const int8_t Q_FIRST[3] = {45, 40, 32};
static const int8_t Q_THRESH_MULTS[8] = {32, 25, 19, 19, 0, 18, 0, 17};
static const int8_t W_MULT[8] = {85, 51, 37, 28, 0, 20, 0, 15};

/* quantizer.rs:31-47 */
int dq_lookup(int qidx) {
    qidx = iclip(qidx, 0, 255);
    if (qidx == 0) return 64;
    qidx -= 1;
    int shift = qidx / 24;
    return (int)R2T_DQ_LOOKUP[qidx % 24] << shift;
}

/* lf_mask.rs:332-335 */
uint32_t deblock_quant_thr(int hbd, int qidx) {
    int qmax = 255 + 48 * hbd;
    return (uint32_t)((dq_lookup(iclip(qidx, 0, qmax)) + 4) >> (3 + 6));
}

/* lf_mask.rs:337-342 */
uint32_t deblock_side_thr(int hbd, int qidx) {
    int bdm8 = 2 * hbd;
    int q_ind = iclip(qidx - 24 * bdm8, 0, 296 - 1);
    int side = R2T_DEBLOCK_SIDE_THRESHOLDS[q_ind];
    return (uint32_t)imax(side + ((1 << 4) >> bdm8), 0) >> (5 - bdm8);
}

/* deblock.rs:13-32: one entry per segment (one without segmentation) */
static void init_lut_y(const Rav2dFrameParams *p, int hbd, int dir, int qidx, uint32_t lut[2][16]) {
    memset(lut, 0, sizeof(uint32_t) * 32);
    int qmax = 255 + 48 * hbd;
    int n = p->seg_enabled ? 8 : 1;
    for (int i = 0; i < n; i++) {
        int yac = p->seg_enabled ? iclip(qidx + p->seg_delta_q[i], 0, qmax) : qidx;
        int dir_yac = yac + 8 * p->db_delta_q_y[dir];
        lut[0][i] = deblock_quant_thr(hbd, dir_yac);
        lut[1][i] = deblock_side_thr(hbd, dir_yac);
    }
}

/* deblock.rs:34-57 */
static void init_lut_uv(const Rav2dFrameParams *p, int hbd, int qidx, uint32_t lut[2][2][16]) {
    memset(lut, 0, sizeof(uint32_t) * 64);
    int qmax = 255 + 48 * hbd;
    int n = p->seg_enabled ? 8 : 1;
    for (int i = 0; i < n; i++) {
        int yac = p->seg_enabled ? iclip(qidx + p->seg_delta_q[i], 0, qmax) : qidx;
        int uac = yac + p->uac_delta + 8 * p->db_delta_q_u;
        lut[0][0][i] = deblock_quant_thr(hbd, uac);
        lut[0][1][i] = deblock_side_thr(hbd, uac);
        int vac = yac + p->vac_delta + 8 * p->db_delta_q_v;
        lut[1][0][i] = deblock_quant_thr(hbd, vac);
        lut[1][1][i] = deblock_side_thr(hbd, vac);
    }
}

/* deblock.rs:81-186 */
template <typename P>
static int filter_choice(const P *buf, ptrdiff_t s, ptrdiff_t t, ptrdiff_t stride, int max_width_neg,
                         int max_width_pos, uint32_t q_thr, uint32_t side_thr) {
    auto at = [&](ptrdiff_t off) -> int { return (int)buf[off]; };
    auto uabs = [](int v) -> uint32_t { return (uint32_t)(v < 0 ? -v : v); };
    uint32_t sd[4];
    for (int dist = -2; dist < 2; dist++) {
        ptrdiff_t d = dist;
        uint32_t ds = uabs(at(s + (d - 1) * stride) - at(s + d * stride) * 2 + at(s + (d + 1) * stride));
        uint32_t dt = uabs(at(t + (d - 1) * stride) - at(t + d * stride) * 2 + at(t + (d + 1) * stride));
        sd[dist + 2] = (ds + dt + 1) >> 1;
    }
    uint32_t high = std::max(sd[0], sd[3]);
    if (high > side_thr) return 0;
    if (max_width_pos == 1) return 1;
    uint32_t side_thr2 = side_thr >> 2;
    uint32_t transition = sd[1] + sd[2];
    if (high > side_thr2) return 1;
    if (transition > q_thr * 4) return 1;
    uint32_t side_thr3 = side_thr >> 3;
    if (high > side_thr3) return 2;
    if (transition > q_thr * 3) return 2;
    uint32_t end_thr = (side_thr * 3) >> 4;
    if (max_width_neg >= 3) {
        uint32_t ds = uabs(at(s - stride) - at(s - 4 * stride) - 3 * (at(s - stride) - at(s - 2 * stride)));
        uint32_t dt = uabs(at(t - stride) - at(t - 4 * stride) - 3 * (at(t - stride) - at(t - 2 * stride)));
        if (((ds + dt + 1) >> 1) > end_thr) return 2;
    }
    {
        uint32_t ds = uabs(at(s) - at(s + 3 * stride) - 3 * (at(s) - at(s + stride)));
        uint32_t dt = uabs(at(t) - at(t + 3 * stride) - 3 * (at(t) - at(t + stride)));
        if (((ds + dt + 1) >> 1) > end_thr) return 2;
    }
    if (max_width_pos == 3) return 3;
    transition <<= 4;
    int prev_dist = 3;
    for (int dist = 4; dist <= max_width_pos; dist += 2) {
        uint32_t q_thr4 = q_thr * (uint32_t)Q_FIRST[(dist - 4) >> 1];
        uint32_t end_thr4 = (side_thr * (uint32_t)dist) >> 4;
        if (transition > q_thr4) return prev_dist;
        int dist2 = imin(7, dist);
        if (max_width_neg >= dist2) {
            uint32_t ds = uabs(at(s - stride) - at(s + (-(ptrdiff_t)dist2 - 1) * stride) -
                               dist2 * (at(s - stride) - at(s - 2 * stride)));
            uint32_t dt = uabs(at(t - stride) - at(t + (-(ptrdiff_t)dist2 - 1) * stride) -
                               dist2 * (at(t - stride) - at(t - 2 * stride)));
            if (((ds + dt + 1) >> 1) > end_thr4) return prev_dist;
        }
        uint32_t ds = uabs(at(s) - at(s + dist2 * stride) - dist2 * (at(s) - at(s + stride)));
        uint32_t dt = uabs(at(t) - at(t + dist2 * stride) - dist2 * (at(t) - at(t + stride)));
        if (((ds + dt + 1) >> 1) > end_thr4) return prev_dist;
        prev_dist = dist;
    }
    return max_width_pos;
}

/* deblock.rs:217-283 */
template <typename P>
static void deblock_edge(P *dst, ptrdiff_t off, int bdmax, uint32_t q_thr, uint32_t side_thr,
                         ptrdiff_t stridea, ptrdiff_t strideb, int max_width_pos, int max_width_neg,
                         bool pos_lossless, bool neg_lossless) {
    int width = filter_choice(dst, off, off + 3 * stridea, strideb, max_width_neg, max_width_pos, q_thr,
                              side_thr);
    int width_neg = imin(width, max_width_neg);
    int width_pos = width;
    if (width_pos < 1) return;
    int q_thr_clamp = (int)q_thr * Q_THRESH_MULTS[width - 1];
    ptrdiff_t dp = off;
    for (int i = 0; i < 4; i++) {
        int d0 = dst[dp], dm1 = dst[dp - strideb], dp1 = dst[dp + strideb], dm2 = dst[dp - 2 * strideb];
        int delta_m2 = iclip(4 * (3 * (d0 - dm1) - (dp1 - dm2)), -q_thr_clamp, q_thr_clamp);
        if (!neg_lossless) {
            int dneg = delta_m2 * W_MULT[width_neg - 1];
            for (int j = 0; j < width_neg; j++) {
                ptrdiff_t idx = dp + (-(ptrdiff_t)j - 1) * strideb;
                int diff = (dneg * (width_neg - j) + (1 << 10)) >> 11;
                dst[idx] = (P)iclip((int)dst[idx] + diff, 0, bdmax);
            }
        }
        if (!pos_lossless) {
            int dpos = delta_m2 * W_MULT[width_pos - 1];
            for (int j = 0; j < width_pos; j++) {
                ptrdiff_t idx = dp + j * strideb;
                int diff = (dpos * (width_pos - j) + (1 << 10)) >> 11;
                dst[idx] = (P)iclip((int)dst[idx] - diff, 0, bdmax);
            }
        }
        dp += stridea;
    }
}

/* deblock_{h,v}_sb64{y,uv}_bd, deblock.rs:303-551.  `dir_h` selects the
 * h-variant (walks rows of one column: vertical edges). */
template <typename P>
static void deblock_sb64(P *dst, ptrdiff_t dst_off, ptrdiff_t stride, int bdmax, const uint16_t *vmask,
                         int nmask, const uint16_t *ll_mask, const uint16_t *q_thr, const uint16_t *side_thr,
                         bool edge, bool dir_h, bool chroma) {
    uint32_t vm = 0;
    for (int i = 0; i < nmask; i++) vm |= vmask[i];
    uint32_t y = 1;
    ptrdiff_t dp = dst_off;
    int qi = 0;
    while ((vm & ~(y - 1)) != 0) {
        if (vm & y) {
            int idx;
            if (!chroma) {
                idx = (vmask[3] & y) ? 3 : (vmask[2] & y) ? 2 : ((vmask[1] & y) != 0);
            } else {
                idx = (vmask[2] & y) ? 2 : ((vmask[1] & y) != 0);
            }
            int max_width_pos = chroma ? MAX_WIDTH_UV[idx] : MAX_WIDTH_Y[idx];
            int max_width_neg = edge ? imin(chroma ? 2 : 6, max_width_pos) : max_width_pos;
            deblock_edge(dst, dp, bdmax, q_thr[qi], side_thr[qi], dir_h ? stride : 1, dir_h ? 1 : stride,
                         max_width_pos, max_width_neg, (ll_mask[1] & y) != 0, (ll_mask[0] & y) != 0);
        }
        y <<= 1;
        dp += dir_h ? 4 * stride : 4;
        qi++;
    }
}

static inline int edge_thr(int cur, int prev) { return (cur && prev) ? (cur + prev + 1) >> 1 : (cur | prev); }

/* segment id at seg[off] (0 without a segment map) */
static inline int seg_at(const uint8_t *seg, ptrdiff_t off) { return seg ? seg[off] : 0; }

/* deblock.rs:1477-1516 */
static void setup_thr_cols(uint16_t q_dst[256], uint16_t s_dst[256], const uint8_t *seg, ptrdiff_t seg_off,
                           ptrdiff_t seg_stride, const uint16_t (*mask)[5][4], int bx4_base,
                           const uint32_t lut[2][16], uint16_t left_q[16], uint16_t left_s[16], int y64,
                           int ss_ver, int w4, int h4) {
    int mask_idx = y64 >> ss_ver;
    int mask_shift = (y64 & ss_ver) ? 8 : 0;
    for (int y4 = 0; y4 < h4; y4++) {
        int prev_q = left_q[y4], prev_s = left_s[y4];
        for (int x4 = 0; x4 < w4; x4++) {
            int seg_id = seg_at(seg, seg_off + x4 + y4 * seg_stride);
            int cur_q = (int)lut[0][seg_id], cur_s = (int)lut[1][seg_id];
            int subpu = 3 * ((mask[bx4_base + x4][4][mask_idx] >> (mask_shift + y4)) & 1);
            q_dst[x4 * 16 + y4] = (uint16_t)(edge_thr(cur_q, prev_q) >> subpu);
            s_dst[x4 * 16 + y4] = (uint16_t)(edge_thr(cur_s, prev_s) >> subpu);
            prev_q = cur_q;
            prev_s = cur_s;
        }
        left_q[y4] = (uint16_t)prev_q;
        left_s[y4] = (uint16_t)prev_s;
    }
}

/* deblock.rs:1520-1567 */
static void setup_thr_rows(uint16_t q_dst[256], uint16_t s_dst[256], const uint8_t *seg, ptrdiff_t seg_off,
                           ptrdiff_t seg_stride, const uint16_t (*mask)[5][4], int starty4,
                           const uint32_t lut[2][16], const uint32_t (*above_lut)[16], ptrdiff_t above_off,
                           int sb64x, int ss_hor, int w4, int h4) {
    int mask_idx = sb64x >> ss_hor;
    int mask_shift = (sb64x & ss_hor) ? 8 : 0;
    uint16_t aq[16] = {0}, as_[16] = {0};
    if (above_lut) {
        for (int x4 = 0; x4 < w4; x4++) {
            int seg_id = seg_at(seg, above_off + x4);
            aq[x4] = (uint16_t)above_lut[0][seg_id];
            as_[x4] = (uint16_t)above_lut[1][seg_id];
        }
    }
    for (int x4 = 0; x4 < w4; x4++) {
        int prev_q = aq[x4], prev_s = as_[x4];
        for (int y4 = 0; y4 < h4; y4++) {
            int seg_id = seg_at(seg, seg_off + x4 + y4 * seg_stride);
            int cur_q = (int)lut[0][seg_id], cur_s = (int)lut[1][seg_id];
            int subpu = 3 * ((mask[starty4 + y4][4][mask_idx] >> (mask_shift + x4)) & 1);
            q_dst[x4 + y4 * 16] = (uint16_t)(edge_thr(cur_q, prev_q) >> subpu);
            s_dst[x4 + y4 * 16] = (uint16_t)(edge_thr(cur_s, prev_s) >> subpu);
            prev_q = cur_q;
            prev_s = cur_s;
        }
    }
}

/* lf_mask.rs:344-360 */
static void transpose_ll(uint16_t dst[17], const uint16_t (*src)[4], int x64, int ss_hor, int ss_ver) {
    dst[0] = dst[16 >> ss_hor];
    for (int x = 0; x < (16 >> ss_hor); x++) {
        uint32_t col = 0;
        for (int y = 0; y < (16 >> ss_ver); y++) col |= (uint32_t)(1 & (src[y][x64] >> x)) << y;
        dst[x + 1] = (uint16_t)col;
    }
}

/* deblock.rs:1572-1606 */
static void crop_bottom_edge(Rav2dAv2Filter *mask, size_t n_mask, size_t mask_row, const Geom &g, int sby) {
    int y64_start = sby << g.sb128;
    int y64_end = imin((sby + 1) << g.sb128, (g.bh + 15) >> 4);
    for (int y64 = y64_start; y64 < y64_end; y64++) {
        if ((y64 + 1) * 16 + 4 <= g.bh) continue;
        int starty4 = (y64 * 16) & 0x30;
        int h4 = imin(g.bh - y64 * 16, 16);
        int crop = starty4 + h4 - 2;
        if (crop < 0) continue;
        for (int x256 = 0; x256 < g.sb256w; x256++) {
            if (mask_row + x256 >= n_mask) break;
            int w = imin(64, g.bw - x256 * 64);
            uint16_t (*yv)[4] = mask[mask_row + x256].filter_y[1][crop];
            for (int i = 0; i < ((w + 15) >> 4); i++) {
                uint16_t m = yv[3][i];
                yv[3][i] = 0;
                yv[2][i] |= m;
            }
        }
    }
}

/* deblock.rs:1623-1862 */
template <typename P>
static void deblock64_cols(const Rav2dFrameParams *p, const Geom &g, const Rav2dAv2Filter *mask, size_t n_mask,
                           size_t mask_row, Frame<P> &f, ptrdiff_t y_off, ptrdiff_t uv_off, int y64) {
    int starty4 = (y64 * 16) & 0x30;
    int h4 = imin(g.bh - y64 * 16, 16);
    int uv_h4 = h4 >> g.ss_ver;
    int y64idx = (y64 & 3) << 2;
    int n64 = (g.bw + 15) >> 4;
    if (p->db_level_y[0]) {
        int l_qidx = -1;
        uint32_t lut[2][16] = {{0}};
        uint16_t left_q[16] = {0}, left_s[16] = {0}, ll[17] = {0};
        for (int x64 = 0; x64 < n64; x64++) {
            bool have_left = x64 > 0;
            size_t col = mask_row + (x64 >> 2);
            if (col >= n_mask) break;
            const Rav2dAv2Filter &cl = mask[col];
            int cur_qidx = cl.qidx[(x64 & 3) + y64idx];
            if (cur_qidx != l_qidx) {
                init_lut_y(p, g.hbd, 0, cur_qidx, lut);
                l_qidx = cur_qidx;
            }
            int bx4_base = (x64 & 3) * 16;
            int w4 = imin(g.bw - x64 * 16, 16);
            uint16_t q[256] = {0}, s[256] = {0};
            setup_thr_cols(q, s, g.seg, (ptrdiff_t)y64 * 16 * g.seg_stride + x64 * 16, g.seg_stride,
                           cl.filter_y[0], bx4_base, lut, left_q, left_s, y64 & 3, 0, w4, h4);
            transpose_ll(ll, &cl.lossless_mask_y[starty4], x64 & 3, 0, 0);
            ptrdiff_t cur = y_off + (ptrdiff_t)x64 * 64;
            for (int x = 0; x < w4; x++) {
                if (!have_left && x == 0) continue;
                const uint16_t (*hm)[4] = cl.filter_y[0][bx4_base + x];
                int sb64y = y64 & 3;
                uint16_t vmask[4] = {hm[0][sb64y], hm[1][sb64y], hm[2][sb64y], hm[3][sb64y]};
                uint16_t llm[2] = {ll[x], ll[x + 1]};
                deblock_sb64(f.y, cur + x * 4, f.ys, g.bdmax, vmask, 4, llm, &q[x * 16], &s[x * 16], false,
                             true, false);
            }
        }
    }
    if (p->db_level_u == 0 && p->db_level_v == 0) return;
    if (g.layout == RAV2D_PIXEL_LAYOUT_I400) return;
    int prev_qidx = -1;
    uint32_t lut[2][2][16] = {{{0}}};
    uint16_t left_q[2][16] = {{0}}, left_s[2][16] = {{0}}, ll[17] = {0};
    for (int x64 = 0; x64 < n64; x64++) {
        bool have_left = x64 > 0;
        size_t col = mask_row + (x64 >> 2);
        if (col >= n_mask) break;
        const Rav2dAv2Filter &cl = mask[col];
        int cur_qidx = cl.qidx[(x64 & 3) + y64idx];
        if (cur_qidx != prev_qidx) {
            init_lut_uv(p, g.hbd, cur_qidx, lut);
            prev_qidx = cur_qidx;
        }
        int bx4_base = ((x64 & 3) * 16) >> g.ss_hor;
        int uv_w4 = imin(g.bw - x64 * 16, 16) >> g.ss_hor;
        uint16_t q[2][256] = {{0}}, s[2][256] = {{0}};
        for (int pl = 0; pl < 2; pl++)
            setup_thr_cols(q[pl], s[pl], g.seg_uv,
                           (ptrdiff_t)y64 * (16 >> g.ss_ver) * g.uv_seg_stride + x64 * (16 >> g.ss_hor),
                           g.uv_seg_stride, cl.filter_uv[0], bx4_base, lut[pl], left_q[pl], left_s[pl], y64 & 3,
                           g.ss_ver, uv_w4, uv_h4);
        transpose_ll(ll, &cl.lossless_mask_uv[starty4 >> g.ss_ver], x64 & 3, g.ss_hor, g.ss_ver);
        ptrdiff_t cur = uv_off + (ptrdiff_t)x64 * (64 >> g.ss_hor);
        int mask_idx = (y64 & 3) >> g.ss_ver;
        int mask_shift = ((y64 & 3) & g.ss_ver) ? 8 : 0;
        uint32_t bytes_mask = g.ss_ver ? 0xff : 0xffff;
        for (int x = 0; x < uv_w4; x++) {
            if (!have_left && x == 0) continue;
            const uint16_t (*hm)[4] = cl.filter_uv[0][bx4_base + x];
            uint16_t vmask[3];
            for (int k = 0; k < 3; k++) vmask[k] = (uint16_t)(((uint32_t)hm[k][mask_idx] >> mask_shift) & bytes_mask);
            uint16_t llm[2] = {ll[x], ll[x + 1]};
            if (p->db_level_u)
                deblock_sb64(f.u, cur + x * 4, f.uvs, g.bdmax, vmask, 3, llm, &q[0][x * 16], &s[0][x * 16], false,
                             true, true);
            if (p->db_level_v)
                deblock_sb64(f.v, cur + x * 4, f.uvs, g.bdmax, vmask, 3, llm, &q[1][x * 16], &s[1][x * 16], false,
                             true, true);
        }
    }
}

/* deblock.rs:1866-2140 */
template <typename P>
static void deblock64_rows(const Rav2dFrameParams *p, const Geom &g, const Rav2dAv2Filter *mask, size_t n_mask,
                           size_t mask_row, Frame<P> &f, ptrdiff_t y_off, ptrdiff_t uv_off, int y64) {
    bool have_top = y64 > 0;
    int starty4 = (y64 * 16) & 0x30;
    int h4 = imin(g.bh - y64 * 16, 16);
    int uv_h4 = h4 >> g.ss_ver;
    int y64idx = (y64 & 3) << 2;
    int a_y64idx = ((y64 + 3) & 3) << 2;
    bool has_arow = false;
    size_t a_row = 0;
    if (have_top) {
        size_t above = starty4 == 0 ? (size_t)g.sb256w : 0;
        if (mask_row >= above) {
            has_arow = true;
            a_row = mask_row - above;
        }
    }
    int n64 = (g.bw + 15) >> 4;
    if (p->db_level_y[1]) {
        int l_qidx = -1, al_qidx = -1;
        uint32_t lut[2][16] = {{0}}, a_lut[2][16] = {{0}};
        uint16_t ll[17] = {0};
        for (int x64 = 0; x64 < n64; x64++) {
            size_t col = mask_row + (x64 >> 2);
            if (col >= n_mask) break;
            const Rav2dAv2Filter &cl = mask[col];
            for (int y = 0; y < h4; y++) ll[y + 1] = cl.lossless_mask_y[starty4 + y][x64 & 3];
            int cur_qidx = cl.qidx[(x64 & 3) + y64idx];
            if (cur_qidx != l_qidx) {
                init_lut_y(p, g.hbd, 1, cur_qidx, lut);
                l_qidx = cur_qidx;
            }
            const uint32_t (*above)[16] = nullptr;
            if (has_arow) {
                size_t acol = a_row + (x64 >> 2);
                if (acol < n_mask) {
                    const Rav2dAv2Filter &al = mask[acol];
                    ll[0] = al.lossless_mask_y[(starty4 + 63) & 63][x64 & 3];
                    int a_qidx = al.qidx[(x64 & 3) + a_y64idx];
                    if (a_qidx != al_qidx) {
                        init_lut_y(p, g.hbd, 1, a_qidx, a_lut);
                        al_qidx = a_qidx;
                    }
                    above = a_lut;
                }
            }
            int w4 = imin(g.bw - x64 * 16, 16);
            uint16_t q[256] = {0}, s[256] = {0};
            ptrdiff_t seg_off = (ptrdiff_t)y64 * 16 * g.seg_stride + x64 * 16;
            setup_thr_rows(q, s, g.seg, seg_off, g.seg_stride, cl.filter_y[1], starty4, lut, above,
                           seg_off - g.seg_stride, x64 & 3, 0, w4, h4);
            ptrdiff_t cur = y_off + (ptrdiff_t)x64 * 64;
            for (int y = 0; y < h4; y++) {
                if (!have_top && y == 0) continue;
                const uint16_t (*row)[4] = cl.filter_y[1][starty4 + y];
                uint16_t vmask[4] = {row[0][x64 & 3], row[1][x64 & 3], row[2][x64 & 3], row[3][x64 & 3]};
                uint16_t llm[2] = {ll[y], ll[y + 1]};
                deblock_sb64(f.y, cur + (ptrdiff_t)y * 4 * f.ys, f.ys, g.bdmax, vmask, 4, llm, &q[y * 16],
                             &s[y * 16], y == 0, false, false);
            }
        }
    }
    if (p->db_level_u == 0 && p->db_level_v == 0) return;
    if (g.layout == RAV2D_PIXEL_LAYOUT_I400) return;
    int l_qidx = -1, al_qidx = -1;
    uint32_t lut[2][2][16] = {{{0}}}, a_lut[2][2][16] = {{{0}}};
    uint16_t ll[17] = {0};
    for (int x64 = 0; x64 < n64; x64++) {
        size_t col = mask_row + (x64 >> 2);
        if (col >= n_mask) break;
        const Rav2dAv2Filter &cl = mask[col];
        for (int y = 0; y < uv_h4; y++) ll[y + 1] = cl.lossless_mask_uv[(starty4 >> g.ss_ver) + y][x64 & 3];
        int cur_qidx = cl.qidx[(x64 & 3) + y64idx];
        if (cur_qidx != l_qidx) {
            init_lut_uv(p, g.hbd, cur_qidx, lut);
            l_qidx = cur_qidx;
        }
        bool above_present = false;
        if (has_arow) {
            size_t acol = a_row + (x64 >> 2);
            if (acol < n_mask) {
                const Rav2dAv2Filter &al = mask[acol];
                ll[0] = al.lossless_mask_uv[((starty4 + 63) & 63) >> g.ss_ver][x64 & 3];
                int a_qidx = al.qidx[(x64 & 3) + a_y64idx];
                if (a_qidx != al_qidx) {
                    init_lut_uv(p, g.hbd, a_qidx, a_lut);
                    al_qidx = a_qidx;
                }
                above_present = true;
            }
        }
        int uv_w4 = imin(g.bw - x64 * 16, 16) >> g.ss_hor;
        uint16_t q[2][256] = {{0}}, s[2][256] = {{0}};
        ptrdiff_t uv_seg_off = (ptrdiff_t)y64 * (16 >> g.ss_ver) * g.uv_seg_stride + x64 * (16 >> g.ss_hor);
        for (int pl = 0; pl < 2; pl++)
            setup_thr_rows(q[pl], s[pl], g.seg_uv, uv_seg_off, g.uv_seg_stride, cl.filter_uv[1],
                           starty4 >> g.ss_ver, lut[pl], above_present ? a_lut[pl] : nullptr,
                           uv_seg_off - g.uv_seg_stride, x64 & 3, g.ss_hor, uv_w4, uv_h4);
        ptrdiff_t cur = uv_off + (ptrdiff_t)x64 * (64 >> g.ss_hor);
        int mask_idx = (x64 & 3) >> g.ss_hor;
        int mask_shift = ((x64 & 3) & g.ss_hor) ? 8 : 0;
        uint32_t bytes_mask = g.ss_hor ? 0xff : 0xffff;
        for (int y = 0; y < uv_h4; y++) {
            if (!have_top && y == 0) continue;
            const uint16_t (*row)[4] = cl.filter_uv[1][(starty4 >> g.ss_ver) + y];
            uint16_t vmask[3];
            for (int k = 0; k < 3; k++) vmask[k] = (uint16_t)(((uint32_t)row[k][mask_idx] >> mask_shift) & bytes_mask);
            uint16_t llm[2] = {ll[y], ll[y + 1]};
            if (p->db_level_u)
                deblock_sb64(f.u, cur + (ptrdiff_t)y * 4 * f.uvs, f.uvs, g.bdmax, vmask, 3, llm, &q[0][y * 16],
                             &s[0][y * 16], y == 0, false, true);
            if (p->db_level_v)
                deblock_sb64(f.v, cur + (ptrdiff_t)y * 4 * f.uvs, f.uvs, g.bdmax, vmask, 3, llm, &q[1][y * 16],
                             &s[1][y * 16], y == 0, false, true);
        }
    }
}

/* decode.rs:3744-3810 + deblock.rs:2145-2190.  `mask` is the library-owned
 * copy: the bottom crop mutates it, as in the reference. */
template <typename P>
void deblock_sbrow_impl(const Rav2dFrameParams *p, const Geom &g, Rav2dAv2Filter *mask, size_t n_mask,
                        Frame<P> &f, int sby, bool cols_only) {
    size_t mask_row = (size_t)((sby >> (2 - g.sb128)) * g.sb256w);
    crop_bottom_edge(mask, n_mask, mask_row, g, sby);
    ptrdiff_t y_off0 = (ptrdiff_t)(sby * g.sb_step * 4) * f.ys;
    ptrdiff_t uv_off0 = ((ptrdiff_t)(sby * g.sb_step * 4) * f.uvs) >> g.ss_ver;
    int y64_start = sby << g.sb128;
    int y64_end = imin((sby + 1) << g.sb128, (g.bh + 15) >> 4);
    ptrdiff_t yo = y_off0, uvo = uv_off0;
    for (int y64 = y64_start; y64 < y64_end; y64++) {
        deblock64_cols(p, g, mask, n_mask, mask_row, f, yo, uvo, y64);
        yo += 64 * f.ys;
        uvo += (64 * f.uvs) >> g.ss_ver;
    }
    if (cols_only) return;
    yo = y_off0;
    uvo = uv_off0;
    for (int y64 = y64_start; y64 < y64_end; y64++) {
        deblock64_rows(p, g, mask, n_mask, mask_row, f, yo, uvo, y64);
        yo += 64 * f.ys;
        uvo += (64 * f.uvs) >> g.ss_ver;
    }
}

template void deblock_sbrow_impl<uint8_t>(const Rav2dFrameParams *, const Geom &, Rav2dAv2Filter *, size_t,
                                          Frame<uint8_t> &, int, bool);
template void deblock_sbrow_impl<uint16_t>(const Rav2dFrameParams *, const Geom &, Rav2dAv2Filter *, size_t,
                                           Frame<uint16_t> &, int, bool);

} // namespace oracle

extern "C" int oracle_dq_lookup(int qidx) { return oracle::dq_lookup(qidx); }
extern "C" uint32_t oracle_deblock_quant_thr(int hbd, int qidx) { return oracle::deblock_quant_thr(hbd, qidx); }
extern "C" uint32_t oracle_deblock_side_thr(int hbd, int qidx) { return oracle::deblock_side_thr(hbd, qidx); }

/* Known-answer hooks (deblock.rs:81-186, 1466-1472). */
extern "C" int oracle_filter_choice(const void *buf, int hbd, ptrdiff_t s, ptrdiff_t t, ptrdiff_t stride,
                                    int max_width_neg, int max_width_pos, uint32_t q_thr, uint32_t side_thr) {
    if (hbd)
        return oracle::filter_choice((const uint16_t *)buf, s, t, stride, max_width_neg, max_width_pos, q_thr,
                                     side_thr);
    return oracle::filter_choice((const uint8_t *)buf, s, t, stride, max_width_neg, max_width_pos, q_thr, side_thr);
}
extern "C" int oracle_edge_thr(int cur, int prev) { return oracle::edge_thr(cur, prev); }


/* Segment-level hook for the pinning tests (16-bit storage): deblock_bd on the
 * 4-line segment whose first positive-side sample is dst[off]; stridea steps
 * along the edge, strideb across it. */
extern "C" void oracle_deblock_edge_16(uint16_t *dst, ptrdiff_t off, int bdmax, uint32_t q_thr, uint32_t side_thr,
                                       ptrdiff_t stridea, ptrdiff_t strideb, int max_width_pos, int max_width_neg,
                                       int pos_lossless, int neg_lossless) {
    oracle::deblock_edge(dst, off, bdmax, q_thr, side_thr, stridea, strideb, max_width_pos, max_width_neg,
                         pos_lossless != 0, neg_lossless != 0);
}
/* deblock_{h,v}_sb64{y,uv}_bd (16-bit storage, stride in samples): dir_h = the
 * vertical-edge ("h") walker. */
extern "C" void oracle_deblock_sb64_16(uint16_t *dst, ptrdiff_t stride, int bdmax, const uint16_t *vmask,
                                       const uint16_t *ll_mask, const uint16_t *q_thr, const uint16_t *side_thr,
                                       int edge, int dir_h, int chroma) {
    oracle::deblock_sb64(dst, 0, stride, bdmax, vmask, chroma ? 3 : 4, ll_mask, q_thr, side_thr, edge != 0,
                         dir_h != 0, chroma != 0);
}
