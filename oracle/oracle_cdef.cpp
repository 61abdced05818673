/*
 * Oracle: CDEF with CCSO folded in, restated from crates/rav2d/src/cdef.rs and
 * crates/rav2d/src/ccso.rs, driven per superblock row with the cross-row seam
 * as in filter_sbrow (crates/rav2d/src/decode.rs:3855-3997).
 * TEST INFRASTRUCTURE ONLY (see oracle_common.h).
 */
#include "oracle_cdef.h"
#include "../rav2d_amd/csrc/tables.h"

namespace oracle {

static const int CDEF_HAVE_LEFT = 1, CDEF_HAVE_RIGHT = 2, CDEF_HAVE_TOP = 4, CDEF_HAVE_BOTTOM = 8;
static const int BACKUP_2X8_Y = 1, BACKUP_2X8_UV = 2;

/* ccso.rs:5-13 */
static const int8_t CCSO_POS[7][2] = {{-1, 0}, {0, -1}, {-1, -1}, {-1, 1}, {-1, -2}, {1, -2}, {0, 2}};

/* cdef.rs:117-120 */
static inline int constrain(int diff, int threshold, int shift) {
    int adiff = diff < 0 ? -diff : diff;
    return apply_sign(imin(adiff, imax(0, threshold - (adiff >> shift))), diff);
}

/* cdef.rs:226-236 */
int cdef_adjust_strength(int strength, uint32_t var) {
    if (var == 0) return 0;
    int i = (var >> 6) ? imin(ulog2(var >> 6), 12) : 0;
    return (strength * (4 + i) + 8) >> 4;
}

/* cdef.rs:134-216 (wrapping u32 cost arithmetic) */
template <typename P> int cdef_find_dir(const P *img, ptrdiff_t stride, uint32_t *var, int bdm8) {
    int hv[2][8] = {{0}}, diag[2][15] = {{0}}, alt[4][11] = {{0}};
    for (int y = 0; y < 8; y++)
        for (int x = 0; x < 8; x++) {
            int px = ((int)img[y * stride + x] >> bdm8) - 128;
            diag[0][y + x] += px;
            alt[0][y + (x >> 1)] += px;
            hv[0][y] += px;
            alt[1][3 + y - (x >> 1)] += px;
            diag[1][7 + y - x] += px;
            alt[2][3 - (y >> 1) + x] += px;
            hv[1][x] += px;
            alt[3][(y >> 1) + x] += px;
        }
    auto sq = [](int v) -> uint32_t { return (uint32_t)v * (uint32_t)v; };
    uint32_t cost[8] = {0};
    for (int n = 0; n < 8; n++) {
        cost[2] += sq(hv[0][n]);
        cost[6] += sq(hv[1][n]);
    }
    cost[2] *= 105;
    cost[6] *= 105;
    static const uint32_t DIV[7] = {840, 420, 280, 210, 168, 140, 120};
    for (int n = 0; n < 7; n++) {
        cost[0] += (sq(diag[0][n]) + sq(diag[0][14 - n])) * DIV[n];
        cost[4] += (sq(diag[1][n]) + sq(diag[1][14 - n])) * DIV[n];
    }
    cost[0] += sq(diag[0][7]) * 105;
    cost[4] += sq(diag[1][7]) * 105;
    for (int n = 0; n < 4; n++) {
        int ci = n * 2 + 1;
        for (int m = 0; m < 5; m++) cost[ci] += sq(alt[n][3 + m]);
        cost[ci] *= 105;
        for (int m = 0; m < 3; m++) cost[ci] += (sq(alt[n][m]) + sq(alt[n][10 - m])) * DIV[2 * m + 1];
    }
    int best_dir = 0;
    uint32_t best = cost[0];
    for (int n = 1; n < 8; n++)
        if (cost[n] > best) {
            best = cost[n];
            best_dir = n;
        }
    *var = (best - cost[best_dir ^ 4]) >> 10;
    return best_dir;
}

/* cdef.rs:37-114 + 122-128 */
template <typename P>
static void cdef_padding(int16_t *tmp, int ts, const P *src, ptrdiff_t ss, ptrdiff_t src_off, const P (*left)[2],
                         const P *top, ptrdiff_t top_off, const P *bottom, ptrdiff_t bot_off, int w, int h, int edges) {
    int o = 2 * ts + 2;
    int x_start = -2, x_end = w + 2, y_start = -2, y_end = h + 2;
    auto fill = [&](int base, int fw, int fh) {
        for (int y = 0; y < fh; y++)
            for (int x = 0; x < fw; x++) tmp[base + y * ts + x] = INT16_MIN;
    };
    if (!(edges & CDEF_HAVE_TOP)) {
        fill(o - 2 - 2 * ts, w + 4, 2);
        y_start = 0;
    }
    if (!(edges & CDEF_HAVE_BOTTOM)) {
        fill(o + h * ts - 2, w + 4, 2);
        y_end -= 2;
    }
    if (!(edges & CDEF_HAVE_LEFT)) {
        fill(o + y_start * ts - 2, 2, y_end - y_start);
        x_start = 0;
    }
    if (!(edges & CDEF_HAVE_RIGHT)) {
        fill(o + y_start * ts + w, 2, y_end - y_start);
        x_end -= 2;
    }
    ptrdiff_t toff = top_off;
    for (int y = y_start; y < 0; y++) {
        for (int x = x_start; x < x_end; x++) tmp[o + x + y * ts] = (int16_t)top[toff + x];
        toff += ss;
    }
    for (int y = 0; y < h; y++)
        for (int x = x_start; x < 0; x++) tmp[o + x + y * ts] = (int16_t)left[y][2 + x];
    ptrdiff_t soff = src_off;
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < x_end; x++) tmp[o + x + y * ts] = (int16_t)src[soff + x];
        soff += ss;
    }
    ptrdiff_t boff = bot_off;
    for (int y = h; y < y_end; y++) {
        for (int x = x_start; x < x_end; x++) tmp[o + x + y * ts] = (int16_t)bottom[boff + x];
        boff += ss;
    }
}

/* cdef.rs:320-460 */
template <typename P>
static void cdef_filter_block(P *dst, ptrdiff_t ds, ptrdiff_t dst_off, const P (*left)[2], const P *top,
                              ptrdiff_t top_off, const P *bottom, ptrdiff_t bot_off, int pri, int sec, int dir,
                              int damping, int w, int h, int edges, int bdm8) {
    const int ts = 12;
    int16_t tmp[144];
    int o = 2 * ts + 2;
    cdef_padding(tmp, ts, dst, ds, dst_off, left, top, top_off, bottom, bot_off, w, h, edges);
    ptrdiff_t dp = dst_off;
    int tp = o;
    const int8_t(*D)[2] = R2T_CDEF_DIRECTIONS;
    if (pri) {
        int pri_tap = 4 - ((pri >> bdm8) & 1);
        int pri_shift = imax(0, damping - ulog2((uint32_t)pri));
        if (sec) {
            int sec_shift = damping - ulog2((uint32_t)sec);
            for (int y = 0; y < h; y++) {
                for (int x = 0; x < w; x++) {
                    int px = dst[dp + x];
                    int sum = 0, mx = px, mn = px;
                    int tk = pri_tap;
                    for (int k = 0; k < 2; k++) {
                        int off1 = D[dir + 2][k];
                        int p0 = tmp[tp + x + off1], p1 = tmp[tp + x - off1];
                        sum += tk * constrain(p0 - px, pri, pri_shift);
                        sum += tk * constrain(p1 - px, pri, pri_shift);
                        tk = (tk & 3) | 2;
                        mn = imin(p0, mn);
                        mx = imax(p0, mx);
                        mn = imin(p1, mn);
                        mx = imax(p1, mx);
                        int off2 = D[dir + 4][k], off3 = D[dir][k];
                        int s0 = tmp[tp + x + off2], s1 = tmp[tp + x - off2];
                        int s2 = tmp[tp + x + off3], s3 = tmp[tp + x - off3];
                        int st = 2 - k;
                        sum += st * constrain(s0 - px, sec, sec_shift);
                        sum += st * constrain(s1 - px, sec, sec_shift);
                        sum += st * constrain(s2 - px, sec, sec_shift);
                        sum += st * constrain(s3 - px, sec, sec_shift);
                        mn = imin(s0, mn);
                        mx = imax(s0, mx);
                        mn = imin(s1, mn);
                        mx = imax(s1, mx);
                        mn = imin(s2, mn);
                        mx = imax(s2, mx);
                        mn = imin(s3, mn);
                        mx = imax(s3, mx);
                    }
                    dst[dp + x] = (P)iclip(px + ((sum - (sum < 0) + 8) >> 4), mn, mx);
                }
                dp += ds;
                tp += ts;
            }
        } else {
            for (int y = 0; y < h; y++) {
                for (int x = 0; x < w; x++) {
                    int px = dst[dp + x];
                    int sum = 0, tk = pri_tap;
                    for (int k = 0; k < 2; k++) {
                        int off = D[dir + 2][k];
                        int p0 = tmp[tp + x + off], p1 = tmp[tp + x - off];
                        sum += tk * constrain(p0 - px, pri, pri_shift);
                        sum += tk * constrain(p1 - px, pri, pri_shift);
                        tk = (tk & 3) | 2;
                    }
                    dst[dp + x] = (P)(px + ((sum - (sum < 0) + 8) >> 4));
                }
                dp += ds;
                tp += ts;
            }
        }
    } else {
        int sec_shift = damping - ulog2((uint32_t)sec);
        for (int y = 0; y < h; y++) {
            for (int x = 0; x < w; x++) {
                int px = dst[dp + x];
                int sum = 0;
                for (int k = 0; k < 2; k++) {
                    int off1 = D[dir + 4][k], off2 = D[dir][k];
                    int s0 = tmp[tp + x + off1], s1 = tmp[tp + x - off1];
                    int s2 = tmp[tp + x + off2], s3 = tmp[tp + x - off2];
                    int st = 2 - k;
                    sum += st * constrain(s0 - px, sec, sec_shift);
                    sum += st * constrain(s1 - px, sec, sec_shift);
                    sum += st * constrain(s2 - px, sec, sec_shift);
                    sum += st * constrain(s3 - px, sec, sec_shift);
                }
                dst[dp + x] = (P)(px + ((sum - (sum < 0) + 8) >> 4));
            }
            dp += ds;
            tp += ts;
        }
    }
}

/* ccso.rs:16-24 */
static inline uint32_t ccso_score(int diff, int quant_step, uint32_t edge_clf) {
    if (diff > quant_step && edge_clf == 0) return 2;
    if (diff < -quant_step) return 0;
    return 1;
}

/* ccso.rs:26-64, 115-186 */
template <typename P>
static void ccso_prep(uint8_t *dst, int dst_stride, const P *src, ptrdiff_t ss, ptrdiff_t src_off, const P (*left)[2],
                      const P *top, ptrdiff_t top_off, const P *bottom, ptrdiff_t bot_off, uint32_t max_band_log2,
                      int ext, int quant_step, uint32_t edge_clf, bool bo_only, int edges, int w, int h, int ss_hor,
                      int ss_ver, int bitdepth) {
    uint32_t shift = (uint32_t)bitdepth - max_band_log2;
    int dy = CCSO_POS[ext][0], dx = CCSO_POS[ext][1];
    const int ts = 68;
    int luma_offset = dx + dy * ts;
    std::vector<P> tmp((size_t)ts * (imax(h, 8) * (1 << ss_ver) + 8));
    int o = 2 * ts + 2;
    int pw = w << ss_hor, ph = h << ss_ver;
    int x_min = (edges & CDEF_HAVE_LEFT) ? -2 : 0;
    int x_max = pw - 1 + ((edges & CDEF_HAVE_RIGHT) ? 2 : 0);
    int y_min = (edges & CDEF_HAVE_TOP) ? -2 : 0;
    int y_max = ph - 1 + ((edges & CDEF_HAVE_BOTTOM) ? 2 : 0);
    for (int y = -2; y < ph + 2; y++) {
        int sy = iclip(y, y_min, y_max);
        for (int x = -2; x < pw + 2; x++) {
            int sx = iclip(x, x_min, x_max);
            P v;
            if (sy < 0)
                v = top[top_off + sx + (2 + sy) * ss];
            else if (sy >= ph)
                v = bottom[bot_off + sx + (sy - ph) * ss];
            else if (sx < 0)
                v = left[sy][2 + sx];
            else
                v = src[src_off + sx + sy * ss];
            tmp[o + x + y * ts] = v;
        }
    }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int xl = x << ss_hor;
            int ti = o + (y << ss_ver) * ts + xl;
            int c = tmp[ti];
            uint8_t band = (uint8_t)((uint32_t)c >> shift);
            if (bo_only) {
                dst[y * dst_stride + x] = band;
            } else {
                uint32_t c0 = ccso_score((int)tmp[ti + luma_offset] - c, quant_step, edge_clf);
                uint32_t c1 = ccso_score((int)tmp[ti - luma_offset] - c, quant_step, edge_clf);
                dst[y * dst_stride + x] = (uint8_t)((c0 << 5) | (c1 << 3)) | band;
            }
        }
}

/* ccso.rs:215-254 */
template <typename P>
static void ccso_add(P *dst, ptrdiff_t ds, const uint8_t *idx, int idx_stride, const uint8_t *offs,
                     const int8_t *lut, int w, int h, const uint16_t (*ll)[4], int bdmax) {
    for (int yy = 0; yy < h; yy += 4) {
        int mi = yy >> 2;
        for (int xx = 0, bx = 0; xx < w; xx += 4, bx++) {
            if (ll[mi][0] & (1 << bx)) continue;
            for (int y = yy; y < yy + 4; y++)
                for (int x = xx; x < xx + 4; x++) {
                    uint8_t i = idx[y * idx_stride + x];
                    int bi = i >> 1, hi = i & 1;
                    uint8_t packed = bi < 64 ? offs[bi] : 0;
                    int oi = 7 & (packed >> (4 * hi));
                    dst[y * ds + x] = (P)iclip((int)dst[y * ds + x] + lut[oi], 0, bdmax);
                }
        }
    }
}

template <typename P>
static void backup2x8(P (*dst)[2], const P *src, ptrdiff_t base, ptrdiff_t stride, ptrdiff_t x_off, int rows) {
    ptrdiff_t off = base;
    for (int y = 0; y < rows; y++) {
        ptrdiff_t s = off + x_off - 2;
        dst[y][0] = src[s];
        dst[y][1] = src[s + 1];
        off += stride;
    }
}

/* cdef.rs:611-645 (positive strides) */
template <typename P>
static void backup2lines_bank(std::vector<P> bank[3], const Frame<P> &f, ptrdiff_t y_off, ptrdiff_t uv_off,
                              int layout) {
    for (ptrdiff_t i = 0; i < 2 * f.ys; i++) bank[0][i] = f.y[y_off + 6 * f.ys + i];
    if (layout == RAV2D_PIXEL_LAYOUT_I400) return;  // cdef.rs:525 (I400)
    // the last two rows of the chroma band: rows 2,3 of 4 (4:2:0), 6,7 of 8
    ptrdiff_t s = uv_off + (layout == RAV2D_PIXEL_LAYOUT_I420 ? 2 : 6) * f.uvs;
    for (ptrdiff_t i = 0; i < 2 * f.uvs; i++) {
        bank[1][i] = f.u[s + i];
        bank[2][i] = f.v[s + i];
    }
}

/* cdef.rs:605: chroma direction from the luma direction, identity except 4:2:2 */
static const uint8_t UV_DIRS[2][8] = {{0, 1, 2, 3, 4, 5, 6, 7}, {7, 0, 2, 4, 5, 6, 6, 6}};

/* cdef.rs:702-1182 */
template <typename P>
void cdef_brow(const CdefBrowParams &p, Frame<P> &f, CdefLines<P> &L, int by_start, int by_end, int bitdepth) {
    const int bdm8 = bitdepth - 8;
    const int bdmax = (1 << bitdepth) - 1;
    const int damping = p.damping + bdm8;
    const ptrdiff_t y_ls = f.ys, uv_ls = f.uvs;
    const int ss_hor = p.layout != RAV2D_PIXEL_LAYOUT_I444, ss_ver = p.layout == RAV2D_PIXEL_LAYOUT_I420;
    const uint8_t *uv_dir = UV_DIRS[p.layout == RAV2D_PIXEL_LAYOUT_I422];
    const int sbsz = 16;
    const int sb64w = (p.bw + sbsz - 1) >> 4;
    ptrdiff_t row_y = (ptrdiff_t)by_start * 4 * y_ls;
    ptrdiff_t row_uv = (((ptrdiff_t)by_start * 4) >> ss_ver) * uv_ls;
    int edge_top = by_start > 0 ? CDEF_HAVE_TOP : 0;
    static thread_local uint8_t lut_idx[3][64 * 8];
    for (int by = by_start; by < by_end; by += 2) {
        int tf = L.toggle;
        int by_idx = (by & 0x3e) >> 1;
        int edges = edge_top | CDEF_HAVE_BOTTOM;
        if (by + 2 >= p.bh) edges &= ~CDEF_HAVE_BOTTOM;
        if (edges & CDEF_HAVE_BOTTOM) backup2lines_bank(L.bank[1 - tf], f, row_y, row_uv, p.layout);
        P lr_bak[2][3][8][2];
        memset(lr_bak, 0, sizeof(lr_bak));
        int bit = 0;
        edges &= ~CDEF_HAVE_LEFT;
        edges |= CDEF_HAVE_RIGHT;
        int prev_flag = 0;
        ptrdiff_t sb_y = row_y, sb_uv = row_uv;
        for (int sbx = 0; sbx < sb64w; sbx++) {
            int sb256x = sbx >> 2;
            int sb64x_idx = sbx & 3;
            int sb64_idx = ((by & 0x30) >> 2) + (sbx & 3);
            const Rav2dAv2Filter *m = p.sb(sb256x);
            int cdef_idx = m ? m->cdef_idx[sb64_idx] : -1;
            // CCSO prep (before CDEF), from pre-CDEF luma
            if (p.ccso_on) {
                uint8_t ccm[3] = {0, 0, 0};
                if (m) memcpy(ccm, m->ccso, 3);
                int flag = ccm[0] | ccm[1] | ccm[2];
                int do_left = flag & ~prev_flag;
                prev_flag |= flag;
                if (do_left && (edges & CDEF_HAVE_LEFT)) {
                    if (do_left & BACKUP_2X8_Y) backup2x8(lr_bak[bit][0], f.y, sb_y, y_ls, 0, 8);
                    if (p.have_chroma && (do_left & BACKUP_2X8_UV)) {
                        backup2x8(lr_bak[bit][1], f.u, sb_uv, uv_ls, 0, 8 >> ss_ver);
                        backup2x8(lr_bak[bit][2], f.v, sb_uv, uv_ls, 0, 8 >> ss_ver);
                    }
                }
                for (int pl = 0; pl < 3; pl++) {
                    // a monochrome stream signals no chroma CCSO (its planes do not exist)
                    if (!ccm[pl] || (pl && !p.have_chroma)) continue;
                    const CcsoCfg &cfg = p.ccso[pl];
                    int phor = pl ? ss_hor : 0, pver = pl ? ss_ver : 0;
                    int sb_edges = edges;
                    if ((sbx + 1) * sbsz >= p.bw) sb_edges &= ~CDEF_HAVE_RIGHT;
                    int w_full = imin(sbsz, p.bw - sbx * sbsz) * 4;
                    if (w_full <= 0) continue;
                    ptrdiff_t top_off = (ptrdiff_t)sbx * sbsz * 4;
                    ptrdiff_t bot_off = sb_y + 8 * y_ls;
                    ccso_prep(lut_idx[pl], 64 >> phor, f.y, y_ls, sb_y, lr_bak[bit][0], L.bank[tf][0].data(), top_off,
                              f.y, bot_off, cfg.max_band_log2, cfg.ext_filter, cfg.quant_step, cfg.edge_clf,
                              cfg.bo_only, sb_edges, w_full >> phor, 8 >> pver, phor, pver, bitdepth);
                }
            }
            bool cdef_active = !(cdef_idx == -1 || !p.cdef_on ||
                                 (p.y_strength[imax(cdef_idx, 0)] == 0 && p.uv_strength[imax(cdef_idx, 0)] == 0));
            if (cdef_active) {
                uint16_t noskip_full = p.on_skip_tx ? 0xffff : (m ? m->noskip_mask[by_idx][sb64x_idx] : 0);
                int y_lvl = p.y_strength[cdef_idx], uv_lvl = p.uv_strength[cdef_idx];
                int flag = (y_lvl != 0) + ((uv_lvl != 0) << 1);
                int y_pri = (y_lvl >> 2) << bdm8;
                int y_sec = y_lvl & 3;
                y_sec += y_sec == 3;
                y_sec <<= bdm8;
                int uv_pri = (uv_lvl >> 2) << bdm8;
                int uv_sec = uv_lvl & 3;
                uv_sec += uv_sec == 3;
                uv_sec <<= bdm8;
                // lossless masks are read only on segmented-lossless frames (cdef.rs:894-911)
                uint16_t y_ll0 = 0, y_ll1 = 0, uv_ll0 = 0, uv_ll1 = 0;
                if (p.any_lossless && m) {
                    int yr = imin(2 * by_idx, 63);
                    y_ll0 = m->lossless_mask_y[yr][sb64x_idx];
                    y_ll1 = m->lossless_mask_y[imin(yr + 1, 63)][sb64x_idx];
                    int uvr = imin((2 * by_idx) >> ss_ver, 63);
                    uv_ll0 = m->lossless_mask_uv[uvr][sb64x_idx];
                    uv_ll1 = m->lossless_mask_uv[imin(uvr + (1 - ss_ver), 63)][sb64x_idx];
                }
                ptrdiff_t b_y = sb_y, b_uv = sb_uv;
                int sb_bx_end = imin((sbx + 1) * sbsz, p.bw);
                for (int bx = sbx * sbsz; bx < sb_bx_end; bx += 2) {
                    if (bx + 2 >= p.bw) edges &= ~CDEF_HAVE_RIGHT;
                    uint16_t bx_mask = (uint16_t)(3u << (bx & 14));
                    bool y_lossless = ((y_ll0 | y_ll1) & bx_mask) != 0;
                    uint16_t uvbx_mask = (uint16_t)((3u >> ss_hor) << ((bx & 14) >> ss_hor));
                    bool uv_lossless = ((uv_ll0 | uv_ll1) & uvbx_mask) != 0;
                    if ((noskip_full & bx_mask) == 0 || (y_lossless && uv_lossless)) {
                        prev_flag = 0;
                        edges |= CDEF_HAVE_LEFT;
                        b_y += 8;
                        b_uv += 8 >> ss_hor;
                        continue;
                    }
                    int do_left = flag & ~prev_flag;
                    prev_flag = flag;
                    if (do_left && (edges & CDEF_HAVE_LEFT)) {
                        if (do_left & BACKUP_2X8_Y) backup2x8(lr_bak[bit][0], f.y, b_y, y_ls, 0, 8);
                        if (p.have_chroma && (do_left & BACKUP_2X8_UV)) {
                            backup2x8(lr_bak[bit][1], f.u, b_uv, uv_ls, 0, 8 >> ss_ver);
                            backup2x8(lr_bak[bit][2], f.v, b_uv, uv_ls, 0, 8 >> ss_ver);
                        }
                    }
                    if (edges & CDEF_HAVE_RIGHT) {
                        int other = 1 - bit;
                        if (flag & BACKUP_2X8_Y) backup2x8(lr_bak[other][0], f.y, b_y, y_ls, 8, 8);
                        if (p.have_chroma && (flag & BACKUP_2X8_UV)) {
                            backup2x8(lr_bak[other][1], f.u, b_uv, uv_ls, 8 >> ss_hor, 8 >> ss_ver);
                            backup2x8(lr_bak[other][2], f.v, b_uv, uv_ls, 8 >> ss_hor, 8 >> ss_ver);
                        }
                    }
                    uint32_t variance = 0;
                    int dir = (y_pri || uv_pri) ? cdef_find_dir(&f.y[b_y], y_ls, &variance, bdm8) : 0;
                    ptrdiff_t top_col = (ptrdiff_t)bx * 4;
                    ptrdiff_t bot_y = b_y + 8 * y_ls;
                    if (y_pri) {
                        int adj = cdef_adjust_strength(y_pri, variance);
                        if ((adj || y_sec) && !y_lossless)
                            cdef_filter_block(f.y, y_ls, b_y, lr_bak[bit][0], L.bank[tf][0].data(), top_col, f.y, bot_y,
                                              adj, y_sec, dir, damping, 8, 8, edges, bdm8);
                    } else if (y_sec && !y_lossless) {
                        cdef_filter_block(f.y, y_ls, b_y, lr_bak[bit][0], L.bank[tf][0].data(), top_col, f.y, bot_y, 0,
                                          y_sec, 0, damping, 8, 8, edges, bdm8);
                    }
                    if (uv_lvl && p.have_chroma && !uv_lossless) {
                        int uvdir = uv_pri ? uv_dir[dir] : 0;
                        int cw = 8 >> ss_hor, ch = 8 >> ss_ver;
                        ptrdiff_t top_col_uv = ((ptrdiff_t)bx * 4) >> ss_hor;
                        ptrdiff_t bot_uv = b_uv + ch * uv_ls;
                        cdef_filter_block(f.u, uv_ls, b_uv, lr_bak[bit][1], L.bank[tf][1].data(), top_col_uv, f.u, bot_uv,
                                          uv_pri, uv_sec, uvdir, damping - 1, cw, ch, edges, bdm8);
                        cdef_filter_block(f.v, uv_ls, b_uv, lr_bak[bit][2], L.bank[tf][2].data(), top_col_uv, f.v, bot_uv,
                                          uv_pri, uv_sec, uvdir, damping - 1, cw, ch, edges, bdm8);
                    }
                    bit ^= 1;
                    edges |= CDEF_HAVE_LEFT;
                    b_y += 8;
                    b_uv += 8 >> ss_hor;
                }
            } else {
                prev_flag = 0;
                edges |= CDEF_HAVE_LEFT;
            }
            // CCSO add on the CDEF output (next_sb)
            if (p.ccso_on) {
                uint8_t ccm[3] = {0, 0, 0};
                if (m) memcpy(ccm, m->ccso, 3);
                int flag = ccm[0] | ((ccm[1] | ccm[2]) << 1);
                int do_right = flag & ~prev_flag;
                if (do_right && (sbx + 1) * sbsz < p.bw) {
                    if (do_right & BACKUP_2X8_Y) backup2x8(lr_bak[bit][0], f.y, sb_y, y_ls, sbsz * 4, 8);
                    if (p.have_chroma && (do_right & BACKUP_2X8_UV)) {
                        backup2x8(lr_bak[bit][1], f.u, sb_uv, uv_ls, (sbsz * 4) >> ss_hor, 8 >> ss_ver);
                        backup2x8(lr_bak[bit][2], f.v, sb_uv, uv_ls, (sbsz * 4) >> ss_hor, 8 >> ss_ver);
                    }
                    prev_flag |= do_right;
                }
                for (int pl = 0; pl < 3; pl++) {
                    if (!ccm[pl] || (pl && !p.have_chroma)) continue;
                    const CcsoCfg &cfg = p.ccso[pl];
                    int phor = pl ? ss_hor : 0, pver = pl ? ss_ver : 0;
                    int w_full = imin(sbsz, p.bw - sbx * sbsz) * 4;
                    P *dst = pl == 0 ? f.y : pl == 1 ? f.u : f.v;
                    ptrdiff_t off = pl == 0 ? sb_y : sb_uv;
                    ptrdiff_t ls = pl == 0 ? y_ls : uv_ls;
                    // rows of the lossless masks for the add (cdef.rs:1117-1150)
                    uint16_t ll[2][4] = {{0}};
                    if (p.any_lossless && m) {
                        const uint16_t (*src)[4] = pl == 0 ? m->lossless_mask_y : m->lossless_mask_uv;
                        int base = (2 * by_idx) >> pver;
                        for (int r = 0; r < 2; r++)
                            if (base + r < 64) ll[r][0] = src[base + r][sb64x_idx];
                    }
                    ccso_add(dst + off, ls, lut_idx[pl], 64 >> phor, cfg.filter_off, R2T_CCSO_OFFSET[cfg.scale_idx],
                             w_full >> phor, 8 >> pver, ll, bdmax);
                }
            }
            sb_y += sbsz * 4;
            sb_uv += (sbsz * 4) >> ss_hor;
        }
        row_y += 8 * y_ls;
        row_uv += (8 * uv_ls) >> ss_ver;
        L.toggle ^= 1;
        edge_top = CDEF_HAVE_TOP;
    }
}

template void cdef_brow<uint8_t>(const CdefBrowParams &, Frame<uint8_t> &, CdefLines<uint8_t> &, int, int, int);
template void cdef_brow<uint16_t>(const CdefBrowParams &, Frame<uint16_t> &, CdefLines<uint16_t> &, int, int, int);
template int cdef_find_dir<uint8_t>(const uint8_t *, ptrdiff_t, uint32_t *, int);
template int cdef_find_dir<uint16_t>(const uint16_t *, ptrdiff_t, uint32_t *, int);

}  // namespace oracle

/* Known-answer hooks (cdef.rs:117-120, 134-216, 226-236; ccso.rs:16-24, 115-186, 215-254). */
extern "C" int oracle_cdef_constrain(int diff, int threshold, int shift) { return oracle::constrain(diff, threshold, shift); }
extern "C" int oracle_cdef_adjust_strength(int strength, uint32_t var) { return oracle::cdef_adjust_strength(strength, var); }
extern "C" int oracle_cdef_find_dir(const void *img, int hbd, ptrdiff_t stride, int bdm8, uint32_t *var) {
    if (hbd) return oracle::cdef_find_dir((const uint16_t *)img, stride, var, bdm8);
    return oracle::cdef_find_dir((const uint8_t *)img, stride, var, bdm8);
}
extern "C" uint32_t oracle_ccso_score(int diff, int quant_step, uint32_t edge_clf) {
    return oracle::ccso_score(diff, quant_step, edge_clf);
}
extern "C" void oracle_ccso_prep_8(uint8_t *dst, int dst_stride, const uint8_t *src, ptrdiff_t ss, ptrdiff_t src_off,
                                   const uint8_t (*left)[2], const uint8_t *top, ptrdiff_t top_off,
                                   const uint8_t *bottom, ptrdiff_t bot_off, uint32_t max_band_log2, int ext,
                                   int quant_step, uint32_t edge_clf, int bo_only, int edges, int w, int h,
                                   int ss_hor, int ss_ver) {
    oracle::ccso_prep(dst, dst_stride, src, ss, src_off, left, top, top_off, bottom, bot_off, max_band_log2, ext,
                      quant_step, edge_clf, bo_only != 0, edges, w, h, ss_hor, ss_ver, 8);
}
extern "C" void oracle_ccso_add_8(uint8_t *dst, ptrdiff_t ds, const uint8_t *idx, int idx_stride,
                                  const uint8_t *offs, const int8_t *lut, int w, int h, const uint16_t (*ll)[4]) {
    oracle::ccso_add(dst, ds, idx, idx_stride, offs, lut, w, h, ll, 255);
}


/* Block-level hook for the pinning tests (16-bit storage): cdef_filter_block
 * on dst (stride ds samples); left[y] = the 2 samples left of row y; top and
 * bottom = 2 rows each (stride ds) whose column 0 is the block's column -2. */
extern "C" void oracle_cdef_filter_block_16(uint16_t *dst, ptrdiff_t ds, const uint16_t (*left)[2],
                                            const uint16_t *top, const uint16_t *bottom, int pri, int sec, int dir,
                                            int damping, int w, int h, int edges, int bdm8) {
    oracle::cdef_filter_block(dst, ds, 0, left, top, 2, bottom, 2, pri, sec, dir, damping, w, h, edges, bdm8);
}
/* ccso_prep / ccso_add (16-bit storage, strides in samples); top / bottom at
 * the block's column 0 with src's stride. */
extern "C" void oracle_ccso_prep_16(uint8_t *dst, int dst_stride, const uint16_t *src, ptrdiff_t ss,
                                    const uint16_t (*left)[2], const uint16_t *top, const uint16_t *bottom,
                                    uint32_t max_band_log2, int ext, int quant_step, uint32_t edge_clf, int bo_only,
                                    int edges, int w, int h, int ss_hor, int ss_ver, int bitdepth) {
    oracle::ccso_prep(dst, dst_stride, src, ss, 0, left, top, 0, bottom, 0, max_band_log2, ext, quant_step, edge_clf,
                      bo_only != 0, edges, w, h, ss_hor, ss_ver, bitdepth);
}
extern "C" void oracle_ccso_add_16(uint16_t *dst, ptrdiff_t ds, const uint8_t *idx, int idx_stride,
                                   const uint8_t *offs, const int8_t *lut, int w, int h, const uint16_t (*ll)[4],
                                   int bdmax) {
    oracle::ccso_add(dst, ds, idx, idx_stride, offs, lut, w, h, ll, bdmax);
}
