/*
 * Oracle: film-grain synthesis on the output copy, restated from
 * crates/rav2d/src/filmgrain.rs (get_random_number, round2, generate_scaling_*,
 * generate_grain_y/uv, sample_lut, fgy_32x32xn, fguv_32x32xn, prep_grain,
 * apply_grain_row, apply_grain_mt with one thread) and the output driver
 * apply_grain_to_picture_mt / picture_has_grain (crates/rav2d/src/decode.rs:
 * 4641-4775).  TEST INFRASTRUCTURE ONLY (see oracle_common.h).
 */
#include "oracle.h"
#include "oracle_common.h"
#include "../rav2d_amd/csrc/tables.h"

namespace oracle {

static const int GW = 82, GH = 73, SGW = 44, SGH = 38;

/* filmgrain.rs:11-16 */
static inline uint32_t get_random_number(uint32_t bits, uint32_t *state) {
    uint32_t r = *state;
    uint32_t bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1;
    *state = (r >> 1) | (bit << 15);
    return (*state >> (16 - bits)) & ((1u << bits) - 1);
}

/* filmgrain.rs:18-20 */
static inline int round2(int x, uint32_t shift) { return (x + ((1 << shift) >> 1)) >> shift; }

/* filmgrain.rs:22-57 */
static void generate_scaling_8bpc(const uint8_t (*pts)[2], int num, uint8_t *scaling) {
    if (num == 0) {
        memset(scaling, 0, 256);
        return;
    }
    for (int i = 0; i < pts[0][0]; i++) scaling[i] = pts[0][1];
    for (int i = 0; i < num - 1; i++) {
        int bx = pts[i][0], by = pts[i][1], ex = pts[i + 1][0], ey = pts[i + 1][1];
        int dx = ex - bx, dy = ey - by;
        if (dx <= 0) continue;
        int delta = dy * ((0x10000 + (dx >> 1)) / dx);
        int d = 0x8000;
        for (int x = 0; x < dx; x++) {
            scaling[bx + x] = (uint8_t)(by + (d >> 16));
            d += delta;
        }
    }
    for (int i = pts[num - 1][0]; i < 256; i++) scaling[i] = pts[num - 1][1];
}

/* filmgrain.rs:59-113 */
static void generate_scaling_hbd(int bitdepth, const uint8_t (*pts)[2], int num, uint8_t *scaling) {
    int shift_x = bitdepth - 8;
    int size = 1 << bitdepth;
    if (num == 0) {
        memset(scaling, 0, size);
        return;
    }
    for (int i = 0; i < (pts[0][0] << shift_x); i++) scaling[i] = pts[0][1];
    for (int i = 0; i < num - 1; i++) {
        int bx = pts[i][0], by = pts[i][1], ex = pts[i + 1][0], ey = pts[i + 1][1];
        int dx = ex - bx, dy = ey - by;
        if (dx <= 0) continue;
        int delta = dy * ((0x10000 + (dx >> 1)) / dx);
        int d = 0x8000;
        for (int x = 0; x < dx; x++) {
            scaling[(bx + x) << shift_x] = (uint8_t)(by + (d >> 16));
            d += delta;
        }
    }
    for (int i = pts[num - 1][0] << shift_x; i < size; i++) scaling[i] = pts[num - 1][1];
    int pad = 1 << shift_x, rnd = pad >> 1;
    for (int i = 0; i < num - 1; i++) {
        int bx = pts[i][0] << shift_x, ex = pts[i + 1][0] << shift_x;
        int dx = ex - bx;
        for (int x = 0; x < dx; x += pad) {
            int base = scaling[bx + x];
            int range = scaling[bx + x + pad] - base;
            int r = rnd;
            for (int n = 1; n < pad; n++) {
                r += range;
                scaling[bx + x + n] = (uint8_t)(base + (r >> shift_x));
            }
        }
    }
}

/* filmgrain.rs:115-158 */
static void generate_grain_y(int16_t (*buf)[GW], const Rav2dFilmGrainData &d, uint32_t seed, int bitdepth) {
    int bdm8 = bitdepth - 8;
    int shift = 4 - bdm8 + d.grain_scale_shift;
    int ctr = 128 << bdm8, gmin = -ctr, gmax = ctr - 1;
    for (int y = 0; y < GH; y++)
        for (int x = 0; x < GW; x++) {
            uint32_t v = get_random_number(11, &seed);
            buf[y][x] = (int16_t)round2(R2T_GAUSSIAN_SEQUENCE[v], (uint32_t)shift);
        }
    const int pad = 3, lag = d.ar_coeff_lag;
    for (int y = pad; y < GH; y++)
        for (int x = pad; x < GW - pad; x++) {
            const int8_t *c = d.ar_coeffs[0];
            int sum = 0, ci = 0;
            for (int dy = y - lag; dy <= y; dy++) {
                int dx_end = dy == y ? x : x + lag + 1;
                for (int dx = x - lag; dx < dx_end; dx++) {
                    if (dy == y && dx == x) break;
                    sum += c[ci] * buf[dy][dx];
                    ci++;
                }
            }
            int g = buf[y][x] + round2(sum, (uint32_t)d.ar_coeff_shift);
            buf[y][x] = (int16_t)iclip(g, gmin, gmax);
        }
}

/* filmgrain.rs:160-229 */
static void generate_grain_uv(int16_t (*buf)[GW], const int16_t (*buf_y)[GW], const Rav2dFilmGrainData &d,
                              uint32_t seed, int uv, int subx, int suby, int bitdepth) {
    seed ^= uv ? 0x49d8 : 0xb524;
    int bdm8 = bitdepth - 8;
    int shift = 4 - bdm8 + d.grain_scale_shift;
    int ctr = 128 << bdm8, gmin = -ctr, gmax = ctr - 1;
    int cw = subx ? SGW : GW, chh = suby ? SGH : GH;
    for (int y = 0; y < chh; y++)
        for (int x = 0; x < cw; x++) {
            uint32_t v = get_random_number(11, &seed);
            buf[y][x] = (int16_t)round2(R2T_GAUSSIAN_SEQUENCE[v], (uint32_t)shift);
        }
    const int pad = 3, lag = d.ar_coeff_lag;
    for (int y = pad; y < chh; y++)
        for (int x = pad; x < cw - pad; x++) {
            const int8_t *c = d.ar_coeffs[1 + uv];
            int sum = 0, ci = 0;
            bool done = false;
            for (int dy = y - lag; dy <= y && !done; dy++) {
                int dx_end = dy == y ? x + 1 : x + lag + 1;
                for (int dx = x - lag; dx < dx_end; dx++) {
                    if (dy == y && dx == x) {
                        if (d.num_points[0] > 0) {
                            int lx = ((x - pad) << subx) + pad, ly = ((y - pad) << suby) + pad;
                            int luma = 0;
                            for (int i = 0; i <= suby; i++)
                                for (int j = 0; j <= subx; j++) luma += buf_y[ly + i][lx + j];
                            luma = round2(luma, (uint32_t)(subx + suby));
                            sum += luma * c[ci];
                        }
                        done = true;
                        break;
                    }
                    sum += c[ci] * buf[dy][dx];
                    ci++;
                }
            }
            int g = buf[y][x] + round2(sum, (uint32_t)d.ar_coeff_shift);
            buf[y][x] = (int16_t)iclip(g, gmin, gmax);
        }
}

/* filmgrain.rs:231-246 */
static inline int sample_lut(const int16_t (*lut)[GW], int bs, const int (*off)[2][2], int subx, int suby, int bx,
                             int by, int x, int y) {
    const int *o = off[bx][by];
    int offx = 3 + (2 >> subx) * (3 + o[1]);
    int offy = 3 + (2 >> suby) * (3 + o[0]);
    return lut[offy + y + (bs >> suby) * by][offx + x + (bs >> subx) * bx];
}

static void row_seeds(uint32_t in_seed, int row_num, int rows, uint32_t seed[2]) {
    for (int i = 0; i < rows; i++) {
        seed[i] = in_seed;
        seed[i] ^= (uint32_t)((((row_num - i) * 37 + 178) & 0xFF)) << 8;
        seed[i] ^= (uint32_t)(((row_num - i) * 173 + 105) & 0xFF);
    }
}

static void next_offsets(int off[2][2][2], int rows, uint32_t seed[2], const Rav2dFilmGrainData &d, bool shift_left) {
    if (shift_left)
        for (int i = 0; i < rows; i++)
            for (int n = 0; n < 2; n++) off[1][i][n] = off[0][i][n];
    for (int i = 0; i < rows; i++)
        for (int n = 0; n < 2; n++) {
            off[0][i][n] = (int)(((uint32_t)(3 - d.block_size) * get_random_number(9, &seed[i])) >> 6);
            for (int k = 0; k < 3; k++) get_random_number(16, &seed[i]);
        }
}

/* filmgrain.rs:248-407 */
template <typename P>
static void fgy(P *dst, const P *src, ptrdiff_t stride, const Rav2dFilmGrainData &d, uint32_t in_seed, int pw,
                const uint8_t *scaling, const int16_t (*lut)[GW], int bh, int row_num, int bitdepth) {
    int rows = 1 + (d.overlap_flag && row_num > 0);
    int bdm8 = bitdepth - 8;
    int ctr = 128 << bdm8, gmin = -ctr, gmax = ctr - 1;
    int bs = 16 << d.block_size;
    int minv = d.clip_to_restricted_range ? 16 << bdm8 : 0;
    int maxv = d.clip_to_restricted_range ? 235 << bdm8 : (1 << bitdepth) - 1;
    uint32_t seed[2] = {0, 0};
    row_seeds(in_seed, row_num, rows, seed);
    int off[2][2][2] = {{{0}}};
    static const int w[2][2] = {{27, 17}, {17, 27}};
    auto put = [&](int x, int y, int bx, int grain) {
        ptrdiff_t si = (ptrdiff_t)y * stride + x + bx;
        int sv = src[si];
        /* the reference indexes the LUT unchecked-in-range (a panic above the coded
         * maximum); clamped here, as in the GPU pass */
        int noise = round2(scaling[imin(sv, (1 << bitdepth) - 1)] * grain, (uint32_t)d.scaling_shift);
        dst[si] = (P)iclip(sv + noise, minv, maxv);
    };
    for (int bx = 0; bx < pw; bx += bs) {
        int bw = imin(bs, pw - bx);
        next_offsets(off, rows, seed, d, d.overlap_flag && bx > 0);
        int ystart = (d.overlap_flag && row_num > 0) ? imin(2, bh) : 0;
        int xstart = (d.overlap_flag && bx > 0) ? imin(2, bw) : 0;
        for (int y = ystart; y < bh; y++) {
            for (int x = xstart; x < bw; x++) put(x, y, bx, sample_lut(lut, bs, off, 0, 0, 0, 0, x, y));
            for (int x = 0; x < xstart; x++) {
                int g = sample_lut(lut, bs, off, 0, 0, 0, 0, x, y);
                int old = sample_lut(lut, bs, off, 0, 0, 1, 0, x, y);
                g = iclip(round2(old * w[x][0] + g * w[x][1], 5), gmin, gmax);
                put(x, y, bx, g);
            }
        }
        for (int y = 0; y < ystart; y++) {
            for (int x = xstart; x < bw; x++) {
                int g = sample_lut(lut, bs, off, 0, 0, 0, 0, x, y);
                int old = sample_lut(lut, bs, off, 0, 0, 0, 1, x, y);
                g = iclip(round2(old * w[y][0] + g * w[y][1], 5), gmin, gmax);
                put(x, y, bx, g);
            }
            for (int x = 0; x < xstart; x++) {
                int top = sample_lut(lut, bs, off, 0, 0, 0, 1, x, y);
                int old = sample_lut(lut, bs, off, 0, 0, 1, 1, x, y);
                top = iclip(round2(old * w[x][0] + top * w[x][1], 5), gmin, gmax);
                int g = sample_lut(lut, bs, off, 0, 0, 0, 0, x, y);
                int old2 = sample_lut(lut, bs, off, 0, 0, 1, 0, x, y);
                g = iclip(round2(old2 * w[x][0] + g * w[x][1], 5), gmin, gmax);
                g = iclip(round2(top * w[y][0] + g * w[y][1], 5), gmin, gmax);
                put(x, y, bx, g);
            }
        }
    }
}

/* filmgrain.rs:409-592 */
template <typename P>
static void fguv(P *dst, const P *src, ptrdiff_t stride, const Rav2dFilmGrainData &d, uint32_t in_seed, int pw,
                 const uint8_t *scaling, const int16_t (*lut)[GW], int bh, int row_num, const P *luma,
                 ptrdiff_t lstride, int uv, bool is_id, int sx, int sy, int bitdepth) {
    int rows = 1 + (d.overlap_flag && row_num > 0);
    int bdm8 = bitdepth - 8;
    int ctr = 128 << bdm8, gmin = -ctr, gmax = ctr - 1;
    int bs = 16 << d.block_size;
    int minv = d.clip_to_restricted_range ? 16 << bdm8 : 0;
    int maxv = d.clip_to_restricted_range ? (is_id ? 235 : 240) << bdm8 : (1 << bitdepth) - 1;
    uint32_t seed[2] = {0, 0};
    row_seeds(in_seed, row_num, rows, seed);
    int off[2][2][2] = {{{0}}};
    static const int w[2][2][2] = {{{27, 17}, {17, 27}}, {{23, 22}, {0, 0}}};
    int bx = 0;
    auto put = [&](int x, int y, int grain) {
        int lx = (bx + x) << sx, ly = y << sy;
        int l = luma[(ptrdiff_t)ly * lstride + lx];
        int avg = sx ? (l + luma[(ptrdiff_t)ly * lstride + lx + 1] + 1) >> 1 : l;
        ptrdiff_t si = (ptrdiff_t)y * stride + bx + x;
        int sv = src[si];
        int val;
        if (!d.chroma_scaling_from_luma) {
            int combined = avg * d.uv_luma_mult[uv] + sv * d.uv_mult[uv];
            val = iclip((combined >> 6) + d.uv_offset[uv] * (1 << bdm8), 0, (1 << bitdepth) - 1);
        } else {
            val = avg;
        }
        int noise = round2(scaling[imin(val, (1 << bitdepth) - 1)] * grain, (uint32_t)d.scaling_shift);
        dst[si] = (P)iclip(sv + noise, minv, maxv);
    };
    for (bx = 0; bx < pw; bx += bs >> sx) {
        int bw = imin(bs >> sx, pw - bx);
        next_offsets(off, rows, seed, d, d.overlap_flag && bx > 0);
        int ystart = (d.overlap_flag && row_num > 0) ? imin(2 >> sy, bh) : 0;
        int xstart = (d.overlap_flag && bx > 0) ? imin(2 >> sx, bw) : 0;
        for (int y = ystart; y < bh; y++) {
            for (int x = xstart; x < bw; x++) put(x, y, sample_lut(lut, bs, off, sx, sy, 0, 0, x, y));
            for (int x = 0; x < xstart; x++) {
                int g = sample_lut(lut, bs, off, sx, sy, 0, 0, x, y);
                int old = sample_lut(lut, bs, off, sx, sy, 1, 0, x, y);
                g = iclip(round2(old * w[sx][x][0] + g * w[sx][x][1], 5), gmin, gmax);
                put(x, y, g);
            }
        }
        for (int y = 0; y < ystart; y++) {
            for (int x = xstart; x < bw; x++) {
                int g = sample_lut(lut, bs, off, sx, sy, 0, 0, x, y);
                int old = sample_lut(lut, bs, off, sx, sy, 0, 1, x, y);
                g = iclip(round2(old * w[sy][y][0] + g * w[sy][y][1], 5), gmin, gmax);
                put(x, y, g);
            }
            for (int x = 0; x < xstart; x++) {
                int top = sample_lut(lut, bs, off, sx, sy, 0, 1, x, y);
                int old = sample_lut(lut, bs, off, sx, sy, 1, 1, x, y);
                top = iclip(round2(old * w[sx][x][0] + top * w[sx][x][1], 5), gmin, gmax);
                int g = sample_lut(lut, bs, off, sx, sy, 0, 0, x, y);
                int old2 = sample_lut(lut, bs, off, sx, sy, 1, 0, x, y);
                g = iclip(round2(old2 * w[sx][x][0] + g * w[sx][x][1], 5), gmin, gmax);
                g = iclip(round2(top * w[sy][y][0] + g * w[sy][y][1], 5), gmin, gmax);
                put(x, y, g);
            }
        }
    }
}

/* filmgrain.rs:617-941 + decode.rs:4672-4775, one thread */
template <typename P>
static void apply_grain(const Rav2dFrameParams *p, P *dy, P *du, P *dv, const P *sy_, const P *su, const P *sv_,
                        ptrdiff_t ys, ptrdiff_t uvs) {
    const Rav2dFilmGrainData &d = p->grain;
    /* I400: the chroma slices are empty, so apply_grain_row skips chroma (filmgrain.rs:745-780) */
    const bool has_chroma = p->layout != RAV2D_PIXEL_LAYOUT_I400;
    const int bitdepth = p->bitdepth, w = p->width, h = p->height;
    const uint32_t seed = p->grain_seed;
    const int ssx = p->layout != RAV2D_PIXEL_LAYOUT_I444, ssy = p->layout == RAV2D_PIXEL_LAYOUT_I420;
    static thread_local int16_t lut_y[GH][GW], lut_u[GH][GW], lut_v[GH][GW];
    memset(lut_u, 0, sizeof(lut_u));
    memset(lut_v, 0, sizeof(lut_v));
    generate_grain_y(lut_y, d, seed, bitdepth);
    if (d.num_points[1] > 0 || d.chroma_scaling_from_luma) generate_grain_uv(lut_u, lut_y, d, seed, 0, ssx, ssy, bitdepth);
    if (d.num_points[2] > 0 || d.chroma_scaling_from_luma) generate_grain_uv(lut_v, lut_y, d, seed, 1, ssx, ssy, bitdepth);
    int ssize = 1 << bitdepth;
    std::vector<uint8_t> scaling[3];
    auto gen = [&](std::vector<uint8_t> &dst, int pl) {
        dst.assign(ssize, 0);
        if (bitdepth == 8)
            generate_scaling_8bpc(d.points[pl], d.num_points[pl], dst.data());
        else
            generate_scaling_hbd(bitdepth, d.points[pl], d.num_points[pl], dst.data());
    };
    if (d.num_points[0] > 0) gen(scaling[0], 0);
    if (!d.chroma_scaling_from_luma)
        for (int uv = 0; uv < 2; uv++)
            if (d.num_points[uv + 1] > 0) gen(scaling[uv + 1], uv + 1);
    int bs = 16 << d.block_size;
    int rows = (h + bs - 1) / bs;
    for (int row = 0; row < rows; row++) {
        int row_start = row * bs;
        if (d.num_points[0] > 0 && !scaling[0].empty()) {
            int bh = imin(h - row_start, bs);
            fgy(dy + row_start * ys, sy_ + row_start * ys, ys, d, seed, w, scaling[0].data(), lut_y, bh, row, bitdepth);
        }
        int ch = (imin(h - row_start, bs) + ssy) >> ssy;
        int cw = (w + ssx) >> ssx;
        ptrdiff_t uv_off = (ptrdiff_t)(row_start >> ssy) * uvs;
        ptrdiff_t l_off = (ptrdiff_t)row_start * ys;
        for (int uv = 0; uv < 2 && has_chroma; uv++) {
            const std::vector<uint8_t> &sc = d.chroma_scaling_from_luma ? scaling[0] : scaling[uv + 1];
            bool has = (d.num_points[uv + 1] > 0 || d.chroma_scaling_from_luma) && !sc.empty();
            if (!has) continue;
            P *dst = uv ? dv : du;
            const P *src = uv ? sv_ : su;
            fguv(dst + uv_off, src + uv_off, uvs, d, seed, cw, sc.data(), uv ? lut_v : lut_u, ch, row, sy_ + l_off, ys,
                 uv, d.mc_identity != 0, ssx, ssy, bitdepth);
        }
    }
}

}  // namespace oracle

using namespace oracle;

/* decode.rs:4641-4651 */
extern "C" int oracle_picture_has_grain(const Rav2dFrameParams *p) {
    if (!p->grain_present) return 0;
    const Rav2dFilmGrainData &d = p->grain;
    return d.num_points[0] != 0 || d.num_points[1] != 0 || d.num_points[2] != 0 ||
           (d.clip_to_restricted_range && d.chroma_scaling_from_luma);
}

extern "C" int oracle_apply_grain(const Rav2dFrameParams *p, const void *src_y, const void *src_u,
                                  const void *src_v, void *dst_y, void *dst_u, void *dst_v, ptrdiff_t y_stride,
                                  ptrdiff_t uv_stride) {
    if (!p || p->layout < RAV2D_PIXEL_LAYOUT_I400 || p->layout > RAV2D_PIXEL_LAYOUT_I444) return RAV2D_EINVAL;
    int bps = p->bitdepth > 8 ? 2 : 1;
    const bool mono = p->layout == RAV2D_PIXEL_LAYOUT_I400;
    const int ssx = p->layout != RAV2D_PIXEL_LAYOUT_I444, ssy = p->layout == RAV2D_PIXEL_LAYOUT_I420;
    int cw = mono ? 0 : (p->width + ssx) >> ssx, ch = mono ? 0 : (p->height + ssy) >> ssy;
    /* the display copy starts as the ungrained picture */
    for (int y = 0; y < p->height; y++)
        memcpy((uint8_t *)dst_y + y * y_stride, (const uint8_t *)src_y + y * y_stride, (size_t)p->width * bps);
    for (int y = 0; y < ch; y++) {
        memcpy((uint8_t *)dst_u + y * uv_stride, (const uint8_t *)src_u + y * uv_stride, (size_t)cw * bps);
        memcpy((uint8_t *)dst_v + y * uv_stride, (const uint8_t *)src_v + y * uv_stride, (size_t)cw * bps);
    }
    if (!oracle_picture_has_grain(p)) return 0;
    if (bps == 2)
        apply_grain<uint16_t>(p, (uint16_t *)dst_y, (uint16_t *)dst_u, (uint16_t *)dst_v, (const uint16_t *)src_y,
                              (const uint16_t *)src_u, (const uint16_t *)src_v, y_stride / 2, uv_stride / 2);
    else
        apply_grain<uint8_t>(p, (uint8_t *)dst_y, (uint8_t *)dst_u, (uint8_t *)dst_v, (const uint8_t *)src_y,
                             (const uint8_t *)src_u, (const uint8_t *)src_v, y_stride, uv_stride);
    return 0;
}

/* Known-answer hooks (filmgrain.rs:11-229). */
extern "C" uint32_t oracle_grain_random(uint32_t bits, uint32_t *state) { return oracle::get_random_number(bits, state); }
extern "C" int oracle_grain_round2(int x, uint32_t shift) { return oracle::round2(x, shift); }
/* points: num (x, y) pairs; scaling: 256 entries (8-bit) or 1 << bitdepth */
extern "C" void oracle_generate_scaling(int bitdepth, const uint8_t (*pts)[2], int num, uint8_t *scaling) {
    if (bitdepth == 8)
        oracle::generate_scaling_8bpc(pts, num, scaling);
    else
        oracle::generate_scaling_hbd(bitdepth, pts, num, scaling);
}
/* buf: 73 rows of 82; uv: 0 = U, 1 = V (buf_y = the luma LUT) */
extern "C" void oracle_generate_grain_y(int16_t *buf, const Rav2dFilmGrainData *d, uint32_t seed, int bitdepth) {
    oracle::generate_grain_y((int16_t(*)[oracle::GW])buf, *d, seed, bitdepth);
}
extern "C" void oracle_generate_grain_uv(int16_t *buf, const int16_t *buf_y, const Rav2dFilmGrainData *d,
                                         uint32_t seed, int uv, int subx, int suby, int bitdepth) {
    oracle::generate_grain_uv((int16_t(*)[oracle::GW])buf, (const int16_t(*)[oracle::GW])buf_y, *d, seed, uv, subx,
                              suby, bitdepth);
}


/* Band-level hooks (16-bit storage; strides in samples) for the per-family
 * entry tests: fgy_32x32xn / fguv_32x32xn on one row band. */
extern "C" void oracle_fgy_16(uint16_t *dst, const uint16_t *src, ptrdiff_t stride, const Rav2dFilmGrainData *d,
                              uint32_t seed, int pw, const uint8_t *scaling, const int16_t *lut, int bh, int row_num,
                              int bitdepth) {
    oracle::fgy(dst, src, stride, *d, seed, pw, scaling, (const int16_t(*)[oracle::GW])lut, bh, row_num, bitdepth);
}
extern "C" void oracle_fguv_16(uint16_t *dst, const uint16_t *src, ptrdiff_t stride, const Rav2dFilmGrainData *d,
                               uint32_t seed, int pw, const uint8_t *scaling, const int16_t *lut, int bh, int row_num,
                               const uint16_t *luma, ptrdiff_t lstride, int uv, int is_id, int sx, int sy,
                               int bitdepth) {
    oracle::fguv(dst, src, stride, *d, seed, pw, scaling, (const int16_t(*)[oracle::GW])lut, bh, row_num, luma,
                 lstride, uv, is_id != 0, sx, sy, bitdepth);
}
