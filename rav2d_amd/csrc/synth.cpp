// Seeded synthetic filter-stage workloads (see synth.h).
#include "synth.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "tables.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {  // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    int range(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }  // inclusive
    bool pct(int p) { return (int)(next() % 100) < p; }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double gauss() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

struct Block {
    int bx, by, bs4, tx4;  // luma 4px units
    bool skip, lossless;
    int dc;                // DC offset (luma units of the coded depth)
    // with R2SConfig.mask_flags: transform partition (TxPartition 0..7), its
    // luma transform size, sub-PU level (subpu_flt_lvl; 3 = none), inter
    int part = 0, tw4 = 0, th4 = 0, subpu = 3;
    bool inter = false;
    int seg = 0;  // segment id (R2S_MASK_SEG)
};

}  // namespace

struct R2SFrame {
    R2SConfig cfg;
    Rav2dFrameParams params;
    int pw[3], ph[3];
    std::vector<uint8_t> planes[3];
    std::vector<uint8_t> padded[3];  // the planes with 8 rows / columns of picture padding
    std::vector<Rav2dAv2Filter> mask;
    std::vector<Rav2dAv2Restoration> lr;
    std::vector<uint8_t> seg[2];  // segment maps: luma / chroma 4x4 units
    int seg_stride[2] = {0, 0};
};

namespace {

inline int ilog2(int v) { return 31 ^ __builtin_clz((unsigned)v); }

// lf_mask.rs:58-86 restated over frame-wide level contexts.
void outer_edge(uint16_t (*masks)[4], int pos4, int len4, int lvl_c, uint8_t *ctx) {
    for (int i = 0; i < len4; i++) {
        int p = pos4 + i;
        int lvl = std::min(lvl_c, (int)ctx[i]);
        masks[lvl][p >> 4] |= (uint16_t)(1u << (p & 15));
    }
    for (int i = 0; i < len4; i++) ctx[i] = (uint8_t)lvl_c;
}

// Edge masks of one block (create_db_mask for an intra block, TxPartition::None)
// with tw4 x th4 transforms.  `dir0[x4]` / `dir1[y4]` are the SB256's
// filter_{y,uv}[0] / [1]; pos are in the plane's 4px units relative to the SB256.
void block_edges(uint16_t (*dir0)[5][4], uint16_t (*dir1)[5][4], int bx4, int by4, int w4, int h4, int tw4, int th4,
                 int lim, uint8_t *actx, uint8_t *lctx) {
    int lw = std::min(lim, ilog2(tw4)), lh = std::min(lim, ilog2(th4));
    outer_edge(dir0[bx4], by4, h4, lw, lctx);
    outer_edge(dir1[by4], bx4, w4, lh, actx);
    for (int x = tw4; x < w4; x += tw4)
        for (int y = 0; y < h4; y++) dir0[bx4 + x][lw][(by4 + y) >> 4] |= (uint16_t)(1u << ((by4 + y) & 15));
    for (int y = th4; y < h4; y += th4)
        for (int x = 0; x < w4; x++) dir1[by4 + y][lh][(bx4 + x) >> 4] |= (uint16_t)(1u << ((bx4 + x) & 15));
}

// Luma transform size (4px units) of TxPartition `part` (None, Split, H, V, H4,
// V4, H5, V5) on a square block of s4 units: the square-block rows of
// TX_PART_TBL (crates/rav2d/src/tables.rs:819-850); 8x8 blocks have None..V.
void part_dims(int s4, int part, int &tw4, int &th4) {
    static const int div_w[8] = {1, 2, 1, 2, 1, 4, 2, 4}, div_h[8] = {1, 2, 2, 1, 4, 1, 4, 2};
    tw4 = std::max(1, s4 / div_w[part]);
    th4 = std::max(1, s4 / div_h[part]);
}

// Rows (or columns) [from, from + n) of one edge line, as mask_inner_edges_v/h
// (lf_mask.rs:88-154) OR them into the four 16-bit words of a level.
void or_span(uint16_t *m, int from, int n) {
    for (int i = 0; i < n; i++) m[(from + i) >> 4] |= (uint16_t)(1u << ((from + i) & 15));
}

enum { kPartH5 = 6, kPartV5 = 7 };

// mask_edges_part (lf_mask.rs:156-254): outer and inner transform edges of one
// block; a / l are the above / left level contexts at the block.
void edges_part(uint16_t (*dir0)[5][4], uint16_t (*dir1)[5][4], int by4, int bx4, int w4, int h4, int part,
                int tw4, int th4, int hlim, int vlim, uint8_t *a, uint8_t *l) {
    const int twl4c = std::min(hlim, ilog2(tw4)), thl4c = std::min(vlim, ilog2(th4));
    if (part < kPartH5) {
        outer_edge(dir0[bx4], by4, h4, twl4c, l);
        outer_edge(dir1[by4], bx4, w4, thl4c, a);
        for (int x = tw4; x < w4; x += tw4) or_span(dir0[bx4 + x][twl4c], by4, h4);
        for (int y = th4; y < h4; y += th4) or_span(dir1[by4 + y][thl4c], bx4, w4);
    } else if (part == kPartH5) {
        // quarter-height rows split in two, the middle half one transform twice as wide
        outer_edge(dir1[by4], bx4, w4, thl4c, a);
        outer_edge(dir0[bx4], by4, std::min(th4, h4), twl4c, l);
        if (h4 > th4) {
            outer_edge(dir0[bx4], by4 + th4, std::min(2 * th4, h4 - th4), std::min(twl4c + 1, hlim), l + th4);
            if (h4 > th4 * 3) outer_edge(dir0[bx4], by4 + th4 * 3, std::min(th4, h4 - 3 * th4), twl4c, l + th4 * 3);
        }
        for (int y = th4; y < h4; y += th4 * 2) or_span(dir1[by4 + y][thl4c], bx4, w4);
        for (int x = tw4; x < w4; x += tw4) {
            or_span(dir0[bx4 + x][twl4c], by4, std::min(th4, h4));
            if (h4 > th4 * 3) or_span(dir0[bx4 + x][twl4c], by4 + th4 * 3, h4 - th4 * 3);
        }
    } else {
        outer_edge(dir0[bx4], by4, h4, twl4c, l);
        outer_edge(dir1[by4], bx4, std::min(tw4, w4), thl4c, a);
        if (w4 > tw4) {
            outer_edge(dir1[by4], bx4 + tw4, std::min(2 * tw4, w4 - tw4), std::min(thl4c + 1, vlim), a + tw4);
            if (w4 > tw4 * 3) outer_edge(dir1[by4], bx4 + tw4 * 3, std::min(tw4, w4 - 3 * tw4), thl4c, a + tw4 * 3);
        }
        for (int x = tw4; x < w4; x += tw4 * 2) or_span(dir0[bx4 + x][twl4c], by4, h4);
        for (int y = th4; y < h4; y += th4) {
            or_span(dir1[by4 + y][thl4c], bx4, std::min(tw4, w4));
            if (w4 > tw4 * 3) or_span(dir1[by4 + y][thl4c], bx4 + tw4 * 3, w4 - tw4 * 3);
        }
    }
}

// mask_subpu_edges (lf_mask.rs:256-330): edges every hsz / vsz units inside
// the block at level twl4c / thl4c; bits new at that level also get the
// level-4 (sub-PU threshold) bit where x (y) & ds_mask is non-zero.
void subpu_edges(uint16_t (*dir0)[5][4], uint16_t (*dir1)[5][4], int by4, int bx4, int w4, int h4, int twl4c,
                 int thl4c, int hsz, int vsz, int ds_mask) {
    auto put = [](uint16_t(*line)[4], int lvl, int from, int n, bool mark) {
        uint16_t add[4] = {0, 0, 0, 0};
        or_span(add, from, n);
        for (int e = 0; e < 4; e++) {
            if (!add[e]) continue;
            const uint16_t had = line[lvl][e];
            line[lvl][e] |= add[e];
            if (mark) line[4][e] |= (uint16_t)(add[e] & ~had);
        }
    };
    if (hsz)
        for (int x = hsz; x < w4; x += hsz) put(dir0[bx4 + x], twl4c, by4, h4, (x & ds_mask) != 0);
    if (vsz)
        for (int y = vsz; y < h4; y += vsz) put(dir1[by4 + y], thl4c, bx4, w4, (y & ds_mask) != 0);
}

// create_db_mask (lf_mask.rs:399-502) for one plane of a block; bx4/by4/w4/h4
// are in the plane's 4px units within the SB256, ssx / ssy the plane's
// subsampling (a chroma plane is one with `chroma` set).
void db_mask(uint16_t (*dir0)[5][4], uint16_t (*dir1)[5][4], const Block &b, int chroma, int ssx, int ssy, int bx4,
             int by4, int w4, int h4, int ds_mask, uint8_t *a, uint8_t *l) {
    const int s = b.subpu, bl2 = ilog2(b.bs4);
    int twl4c, thl4c;
    if (!b.inter || !b.skip) {
        int part = 0, tw4, th4;
        if (b.lossless) {
            tw4 = th4 = 1;
        } else if (chroma) {
            // the largest transform of the chroma block, at most 32 samples a side
            tw4 = std::max(1, std::min(b.bs4 >> ssx, 8));
            th4 = std::max(1, std::min(b.bs4 >> ssy, 8));
        } else {
            part = b.part;
            tw4 = b.tw4;
            th4 = b.th4;
        }
        const int hlim = std::min(std::max(s - ssx, 0), 3 - chroma);
        const int vlim = std::min(std::max(s - ssy, 0), 3 - chroma);
        edges_part(dir0, dir1, by4, bx4, w4, h4, part, tw4, th4, hlim, vlim, a, l);
        twl4c = std::min(s, ilog2(tw4));
        thl4c = std::min(s, ilog2(th4));
    } else {
        const int hlim = b.lossless ? 0 : std::min(std::max(std::min(s, bl2) - ssx, 0), 3 - chroma);
        const int vlim = b.lossless ? 0 : std::min(std::max(std::min(s, bl2) - ssy, 0), 3 - chroma);
        outer_edge(dir0[bx4], by4, h4, hlim, l);
        outer_edge(dir1[by4], bx4, w4, vlim, a);
        twl4c = thl4c = s;
    }
    if (s != 3) {
        const int hl = twl4c - (ssx && twl4c), vl = thl4c - (ssy && thl4c);
        subpu_edges(dir0, dir1, by4, bx4, w4, h4, hl, vl, (1 << s) >> ssx, (1 << s) >> ssy, ds_mask);
    }
}

void partition(Rng &r, int x, int y, int s4, int bw, int bh, std::vector<Block> &out, const R2SConfig &c) {
    if (x >= bw || y >= bh) return;
    bool split = s4 > 16 || (s4 > 2 && r.pct(s4 == 16 ? 60 : s4 == 8 ? 50 : 40));
    if (split) {
        int h = s4 >> 1;
        partition(r, x, y, h, bw, bh, out, c);
        partition(r, x + h, y, h, bw, bh, out, c);
        partition(r, x, y + h, h, bw, bh, out, c);
        partition(r, x + h, y + h, h, bw, bh, out, c);
        return;
    }
    Block b;
    b.bx = x;
    b.by = y;
    b.bs4 = s4;
    b.lossless = r.pct(c.lossless_pct);
    b.tx4 = b.lossless ? 1 : (1 << r.range(0, ilog2(s4)));
    b.skip = r.pct(c.skip_pct);
    b.dc = r.range(-24, 24);
    b.tw4 = b.th4 = b.tx4;
    if (c.mask_flags & R2S_MASK_SUBPU) {
        b.inter = r.pct(60);
        if (b.inter && !b.lossless && r.pct(70)) b.subpu = r.range(0, 2);
    }
    if ((c.mask_flags & R2S_MASK_TXPART) && !b.lossless) {
        b.part = r.range(0, s4 == 2 ? 3 : 7);
        part_dims(s4, b.part, b.tw4, b.th4);
    } else if (c.mask_flags) {
        b.part = 0;  // the block's own transform (lossless: 4x4)
        b.tw4 = b.th4 = b.lossless ? 1 : s4;
    }
    if (c.mask_flags & R2S_MASK_SEG) b.seg = b.lossless ? 7 : r.range(0, 6);
    out.push_back(b);
}

// Rows below and columns right of each plane that a decoder's picture holds
// (reconstruction of the blocks that overhang the frame edge, in the padding
// of its picture allocation, picture.rs:77-125) and the deblocking reads.
constexpr int kPad = 8;

// `pl`: the w x h plane; `pad`: the same plane with kPad rows and columns of
// padding ((w + kPad) samples per row), whose padding continues the plane's
// content (its blocks' DC offsets and noise from the padding's own stream `rp`).
template <typename P>
void fill_plane(Rng &r, Rng &rp, P *pl, P *pad, int w, int h, int ssx, int ssy, const std::vector<Block> &blocks,
                int bd, double sigma) {
    int bdmax = (1 << bd) - 1;
    double gx = r.uniform() * 2.0 - 1.0, gy = r.uniform() * 2.0 - 1.0;
    double base = (0.3 + 0.4 * r.uniform()) * bdmax;
    double amp = 0.35 * bdmax;
    std::vector<int> dc((size_t)w * h, 0);
    int scale = 1 << (bd - 8);
    for (const Block &b : blocks) {
        // DC offset per transform block
        int tw = std::max(4, (b.tw4 * 4) >> ssx), th = std::max(4, (b.th4 * 4) >> ssy);
        int bsw = (b.bs4 * 4) >> ssx, bsh = (b.bs4 * 4) >> ssy;
        int x0 = (b.bx * 4) >> ssx, y0 = (b.by * 4) >> ssy;
        for (int ty = 0; ty < bsh; ty += th)
            for (int tx = 0; tx < bsw; tx += tw) {
                int off = (b.dc + r.range(-6, 6)) * scale;
                for (int yy = y0 + ty; yy < std::min(h, y0 + ty + th); yy++)
                    for (int xx = x0 + tx; xx < std::min(w, x0 + tx + tw); xx++) dc[(size_t)yy * w + xx] = off;
            }
    }
    auto sample = [&](int x, int y, double noise) -> P {
        double v = base + amp * (gx * (x / (double)w - 0.5) + gy * (y / (double)h - 0.5)) +
                   amp * 0.25 * sin(x * 0.013 + y * 0.021) + sigma * noise +
                   dc[(size_t)std::min(y, h - 1) * w + std::min(x, w - 1)];
        int iv = (int)lrint(v);
        return (P)(iv < 0 ? 0 : iv > bdmax ? bdmax : iv);
    };
    const int ps = w + kPad;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) pad[(size_t)y * ps + x] = pl[(size_t)y * w + x] = sample(x, y, r.gauss());
    for (int y = 0; y < h + kPad; y++)
        for (int x = y < h ? w : 0; x < ps; x++) pad[(size_t)y * ps + x] = sample(x, y, rp.gauss());
}

void gen_params(Rng &r, R2SFrame *f) {
    const R2SConfig &c = f->cfg;
    Rav2dFrameParams &p = f->params;
    memset(&p, 0, sizeof(p));
    p.width = c.width;
    p.height = c.height;
    p.bitdepth = c.bitdepth;
    p.layout = RAV2D_PIXEL_LAYOUT_I420;
    p.sb128 = c.sb128;
    uint32_t in = 0;
    if (c.tools & R2S_TOOL_DEBLOCK) in |= RAV2D_INLOOPFILTER_DEBLOCK;
    if (c.tools & R2S_TOOL_CDEF) in |= RAV2D_INLOOPFILTER_CDEF;
    if (c.tools & R2S_TOOL_CCSO) in |= RAV2D_INLOOPFILTER_CCSO;
    if (c.tools & R2S_TOOL_WIENER) in |= RAV2D_INLOOPFILTER_WIENER;
    if (c.tools & R2S_TOOL_GDF) in |= RAV2D_INLOOPFILTER_GDF;
    p.inloop_filters = in;
    p.seq_cdef = 1;
    p.cfl_ds_filter_index = r.range(0, 2);
    p.yac = r.range(1, 255);
    p.uac_delta = r.range(-8, 8);
    p.vac_delta = r.range(-8, 8);
    p.seg_enabled = 0;
    // deblock (obu.rs:1293-1323): on/off bits and delta_q in +-2^(df_par_bits-1)
    p.db_level_y[0] = p.db_level_y[1] = 1;
    p.db_level_u = p.db_level_v = 1;
    int off = 1 << (r.range(2, 5) - 1);
    p.db_delta_q_y[0] = (int8_t)r.range(-off, off - 1);
    p.db_delta_q_y[1] = (int8_t)r.range(-off, off - 1);
    p.db_delta_q_u = (int8_t)r.range(-off, off - 1);
    p.db_delta_q_v = (int8_t)r.range(-off, off - 1);
    // cdef (obu.rs:1347-1366)
    p.cdef_damping = (uint8_t)r.range(3, 6);
    p.cdef_n_strengths = (uint8_t)r.range(1, 8);
    p.cdef_on_skiptx = (uint8_t)r.range(0, 1);
    for (int i = 0; i < p.cdef_n_strengths; i++) {
        // pri 0..15, sec code 0..3 (3 -> 4): "pri<<2 | sec" in 6 bits
        p.cdef_y_strength[i] = (uint8_t)r.range(0, 63);
        p.cdef_uv_strength[i] = (uint8_t)r.range(0, 63);
    }
    // ccso (obu.rs:1696-1736)
    for (int pl = 0; pl < 3; pl++) {
        Rav2dCcsoPlane &cp = p.ccso[pl];
        cp.enabled = (uint8_t)r.pct(80);
        cp.bo_only = (uint8_t)r.pct(30);
        cp.scale_idx = (uint8_t)r.range(0, 3);
        if (cp.bo_only) {
            cp.max_band_log2 = (uint8_t)r.range(0, 7);
        } else {
            cp.quant_idx = (uint8_t)r.range(0, 3);
            cp.ext_filter_support = (uint8_t)r.range(0, 6);
            if (R2T_CCSO_QUANT_SZ[cp.scale_idx][cp.quant_idx] != 0) cp.edge_clf = (uint8_t)r.range(0, 1);
            cp.max_band_log2 = (uint8_t)r.range(0, 3);
        }
        for (int i = 0; i < 64; i++) cp.filter_off[i] = (uint8_t)(r.range(0, 7) | (r.range(0, 7) << 4));
    }
    // restoration (obu.rs:1375-1455)
    for (int pl = 0; pl < 3; pl++) {
        Rav2dRestorationPlane &rp = p.restoration[pl];
        rp.restoration_type = (uint8_t)(pl == 0 ? RAV2D_RESTORATION_SWITCHABLE : RAV2D_RESTORATION_NS_WIENER);
        if (rp.restoration_type >= RAV2D_RESTORATION_NS_WIENER) {
            rp.frame_filters_on = (uint8_t)r.range(0, 1);
            // multi-class frame filters are drawn for luma only
            if (rp.frame_filters_on && pl == 0) {
                int val = r.range(0, 7);
                rp.num_classes_idx = (uint8_t)val;
                rp.num_classes = (uint8_t)(1 + val + std::max(val - 3, 0) + std::max(val - 5, 0) * 2);
            } else {
                rp.num_classes_idx = 0;
                rp.num_classes = 1;
            }
            const int8_t(*range)[2] = pl ? R2T_NS_WIENER_COEF_RANGE_UV : R2T_NS_WIENER_COEF_RANGE_Y;
            int nc = pl ? 18 : 16;
            for (int k = 0; k < 16; k++)
                for (int i = 0; i < nc; i++)
                    rp.filter[k][i] = (int8_t)(range[i][1] + r.range(0, (1 << range[i][0]) - 1));
        }
    }
    p.lr_unit_size[0] = (uint8_t)(c.sb128 ? r.range(7, 9) : r.range(6, 9));
    p.lr_unit_size[1] = (uint8_t)std::max(5, p.lr_unit_size[0] - r.range(0, 1));
    // gdf (obu.rs:1325-1345)
    p.gdf_enabled = 1;
    p.gdf_qp_idx = r.range(0, 5);
    p.gdf_scale = r.range(1, 4);
    p.gdf_ref_dst_idx = r.pct(50) ? 0 : r.range(1, 5);
    // film grain (obu.rs:261-328): lag 3, 8 luma points, chroma from luma
    p.grain_present = (c.tools & R2S_TOOL_GRAIN) ? 1 : 0;
    p.grain_seed = (uint32_t)r.range(0, 65535);
    Rav2dFilmGrainData &g = p.grain;
    g.chroma_scaling_from_luma = 1;
    g.num_points[0] = 8;
    int xs[8];
    for (;;) {
        for (int i = 0; i < 8; i++) xs[i] = r.range(0, 255);
        std::sort(xs, xs + 8);
        bool ok = true;
        for (int i = 1; i < 8; i++) ok = ok && xs[i] > xs[i - 1];
        if (ok) break;
    }
    for (int i = 0; i < 8; i++) {
        g.points[0][i][0] = (uint8_t)xs[i];
        g.points[0][i][1] = (uint8_t)r.range(0, 255);
    }
    g.scaling_shift = r.range(8, 11);
    g.ar_coeff_lag = 3;
    for (int pl = 0; pl < 3; pl++)
        for (int i = 0; i < 25; i++) g.ar_coeffs[pl][i] = (int8_t)r.range(-128, 127);
    g.ar_coeff_shift = (uint64_t)r.range(6, 9);
    g.grain_scale_shift = r.range(0, 3);
    g.overlap_flag = r.range(0, 1);
    g.clip_to_restricted_range = r.range(0, 1);
    g.mc_identity = g.clip_to_restricted_range ? r.range(0, 1) : 0;
    g.block_size = r.range(0, 1);
}

// A monochrome sequence's frame: no chroma planes, so no chroma CCSO,
// restoration or grain.  Applied after every draw, so the luma planes, luma
// side info and luma parameters equal those of the 4:2:0 frame of the seed.
void make_monochrome(R2SFrame *f) {
    Rav2dFrameParams &p = f->params;
    p.layout = RAV2D_PIXEL_LAYOUT_I400;
    for (int pl = 1; pl < 3; pl++) {
        p.ccso[pl].enabled = 0;
        p.restoration[pl].restoration_type = RAV2D_RESTORATION_NONE;
        p.grain.num_points[pl] = 0;
    }
    p.grain.chroma_scaling_from_luma = 0;
    for (Rav2dAv2Filter &m : f->mask) m.ccso[1] = m.ccso[2] = 0;
    for (Rav2dAv2Restoration &lr : f->lr)
        for (int pl = 1; pl < 3; pl++)
            for (int u = 0; u < 16; u++) lr.lr[pl][u].restoration_type = RAV2D_RESTORATION_NONE;
    for (int pl = 1; pl < 3; pl++) {
        f->pw[pl] = f->ph[pl] = 0;
        f->planes[pl].clear();
        f->padded[pl].clear();
    }
}

}  // namespace

extern "C" {

R2SFrame *r2s_frame_create(const R2SConfig *cfg) {
    if (!cfg || cfg->width <= 0 || cfg->height <= 0 || (cfg->width & 7) || (cfg->height & 7) ||
        (cfg->bitdepth != 8 && cfg->bitdepth != 10 && cfg->bitdepth != 12))
        return nullptr;
    R2SFrame *f = new R2SFrame();
    f->cfg = *cfg;
    Rng r(cfg->seed);
    gen_params(r, f);
    const int W = cfg->width, H = cfg->height;
    // chroma subsampling of the format (the I400 frame is made from the 4:2:0 one)
    const int cf = cfg->monochrome ? 0 : cfg->chroma_format;
    const int ssx = cf < 2, ssy = cf == 0;
    f->params.layout = cf == 0 ? RAV2D_PIXEL_LAYOUT_I420 : cf == 1 ? RAV2D_PIXEL_LAYOUT_I422 : RAV2D_PIXEL_LAYOUT_I444;
    f->pw[0] = W;
    f->ph[0] = H;
    f->pw[1] = f->pw[2] = (W + ssx) >> ssx;
    f->ph[1] = f->ph[2] = (H + ssy) >> ssy;
    const int bw = ((W + 7) >> 3) << 1, bh = ((H + 7) >> 3) << 1;
    const int sb256w = (bw + 63) >> 6, sb256h = (bh + 63) >> 6;
    f->mask.assign((size_t)sb256w * sb256h, Rav2dAv2Filter());
    memset(f->mask.data(), 0, f->mask.size() * sizeof(Rav2dAv2Filter));
    f->lr.assign((size_t)sb256w * sb256h, Rav2dAv2Restoration());
    memset(f->lr.data(), 0, f->lr.size() * sizeof(Rav2dAv2Restoration));

    // partition: superblocks of 64 or 128, blocks 8x8..64x64
    std::vector<Block> blocks;
    int sb4 = 16 << cfg->sb128;
    for (int y = 0; y < bh; y += sb4)
        for (int x = 0; x < bw; x += sb4) partition(r, x, y, sb4, bw, bh, blocks, *cfg);

    // edge masks, noskip and lossless masks; a TIP frame of frame_mode 2 sets
    // no sub-PU threshold bits (create_db_mask's ds_subpu_mask = 0)
    const int ds_mask = (cfg->mask_flags & R2S_MASK_SUBPU) && r.pct(25) ? 0 : 15;
    std::vector<uint8_t> actx_y(bw + 64, 3), lctx_y(bh + 64, 3), actx_uv(bw + 64, 2), lctx_uv(bh + 64, 2);
    for (const Block &b : blocks) {
        int sbx = b.bx >> 6, sby = b.by >> 6;
        Rav2dAv2Filter &m = f->mask[(size_t)sby * sb256w + sbx];
        int bx4 = b.bx & 63, by4 = b.by & 63;
        int w4 = std::min(bw - b.bx, b.bs4), h4 = std::min(bh - b.by, b.bs4);
        if (cfg->mask_flags) {
            db_mask(m.filter_y[0], m.filter_y[1], b, 0, 0, 0, bx4, by4, w4, h4, ds_mask, &actx_y[b.bx],
                    &lctx_y[b.by]);
            db_mask(m.filter_uv[0], m.filter_uv[1], b, 1, ssx, ssy, bx4 >> ssx, by4 >> ssy, w4 >> ssx, h4 >> ssy,
                    ds_mask, &actx_uv[b.bx >> ssx], &lctx_uv[b.by >> ssy]);
        } else {
            block_edges(m.filter_y[0], m.filter_y[1], bx4, by4, w4, h4, b.tx4, b.tx4, 3, &actx_y[b.bx],
                        &lctx_y[b.by]);
            // chroma transform = the chroma block, at most 32 samples a side
            int ctw4 = b.lossless ? 1 : std::max(1, std::min(b.bs4 >> ssx, 8));
            int cth4 = b.lossless ? 1 : std::max(1, std::min(b.bs4 >> ssy, 8));
            block_edges(m.filter_uv[0], m.filter_uv[1], bx4 >> ssx, by4 >> ssy, w4 >> ssx, h4 >> ssy, ctw4, cth4, 2,
                        &actx_uv[b.bx >> ssx], &lctx_uv[b.by >> ssy]);
        }
        for (int y = 0; y < h4; y++)
            for (int x = 0; x < w4; x++) {
                int yy = by4 + y, xx = bx4 + x;
                if (!b.skip) {
                    m.noskip_mask[yy >> 1][xx >> 4] |= (uint16_t)(1u << (xx & 15));
                    m.lr_noskip_mask[yy][xx >> 4] |= (uint16_t)(1u << (xx & 15));
                }
                if (b.lossless) m.lossless_mask_y[yy][xx >> 4] |= (uint16_t)(1u << (xx & 15));
            }
        if (b.lossless)
            for (int y = 0; y < (h4 >> ssy); y++)
                for (int x = 0; x < (w4 >> ssx); x++) {
                    // chroma 4px units in the SB256: a word per 64 luma columns
                    int yy = (by4 >> ssy) + y, xx = (bx4 >> ssx) + x;
                    m.lossless_mask_uv[yy][xx >> (4 - ssx)] |= (uint16_t)(1u << (xx & ((16 >> ssx) - 1)));
                }
    }

    // segmentation: per-segment qidx deltas, segment 7 lossless (its qidx is 0,
    // as obu.rs:1270-1280 derives segmentation.lossless) when it is used
    if (cfg->mask_flags & R2S_MASK_SEG) {
        Rav2dFrameParams &sp = f->params;
        sp.seg_enabled = 1;
        for (int i = 0; i < 7; i++) sp.seg_delta_q[i] = (int16_t)r.range(-48, 48);
        sp.seg_delta_q[7] = (int16_t)-sp.yac;
        sp.seg_lossless[7] = cfg->lossless_pct > 0;
        f->seg_stride[0] = bw;
        const int cbw = bw >> ssx;
        f->seg_stride[1] = cbw;
        f->seg[0].assign((size_t)bw * bh, 0);
        f->seg[1].assign((size_t)cbw * (bh >> ssy), 0);
        for (const Block &b : blocks) {
            int w4 = std::min(bw - b.bx, b.bs4), h4 = std::min(bh - b.by, b.bs4);
            for (int y = 0; y < h4; y++)
                for (int x = 0; x < w4; x++) f->seg[0][(size_t)(b.by + y) * bw + b.bx + x] = (uint8_t)b.seg;
            for (int y = 0; y < (h4 >> ssy); y++)
                for (int x = 0; x < (w4 >> ssx); x++)
                    f->seg[1][(size_t)((b.by >> ssy) + y) * cbw + (b.bx >> ssx) + x] = (uint8_t)b.seg;
        }
    }

    // per-64x64 / per-SB256 side info
    const Rav2dFrameParams &p = f->params;
    for (size_t i = 0; i < f->mask.size(); i++) {
        Rav2dAv2Filter &m = f->mask[i];
        for (int k = 0; k < 16; k++) {
            m.qidx[k] = (uint16_t)r.range(1, 255);
            m.cdef_idx[k] = (int8_t)r.range(-1, p.cdef_n_strengths - 1);
            m.gdf[k] = (uint8_t)r.pct(cfg->pc_gdf_pct);
        }
        for (int pl = 0; pl < 3; pl++) m.ccso[pl] = (uint8_t)(p.ccso[pl].enabled ? r.pct(70) : 0);
        Rav2dAv2Restoration &lr = f->lr[i];
        for (int pl = 0; pl < 3; pl++) {
            const Rav2dRestorationPlane &rp = p.restoration[pl];
            const int8_t(*range)[2] = pl ? R2T_NS_WIENER_COEF_RANGE_UV : R2T_NS_WIENER_COEF_RANGE_Y;
            int nc = pl ? 18 : 16;
            for (int u = 0; u < 16; u++) {
                Rav2dRestorationUnit &ru = lr.lr[pl][u];
                if (!r.pct(cfg->pc_gdf_pct)) {
                    ru.restoration_type = RAV2D_RESTORATION_NONE;
                } else if (rp.restoration_type == RAV2D_RESTORATION_SWITCHABLE) {
                    ru.restoration_type =
                        (uint8_t)(r.pct(50) ? RAV2D_RESTORATION_PC_WIENER : RAV2D_RESTORATION_NS_WIENER);
                } else {
                    ru.restoration_type = rp.restoration_type;
                }
                // per-unit filters: classes at the raw [16][32] rows, as the parser writes them
                for (int k = 0; k < 16; k++)
                    for (int t = 0; t < nc; t++)
                        ru.ns_filter[k][t] = (int8_t)(range[t][1] + r.range(0, (1 << range[t][0]) - 1));
            }
        }
    }

    // planes
    int bd = cfg->bitdepth;
    double sigma = 4.0 * (1 << (bd - 8));  // 4 at 8 bit, 16 at 10, 64 at 12
    Rng rp(cfg->seed ^ 0x5ad5ad5ad5ad5adull);  // the padding's own stream: the planes do not depend on it
    const int bps = bd > 8 ? 2 : 1;
    for (int pl = 0; pl < 3; pl++) {
        size_t n = (size_t)f->pw[pl] * f->ph[pl];
        f->planes[pl].assign(n * bps, 0);
        f->padded[pl].assign((size_t)(f->pw[pl] + kPad) * (f->ph[pl] + kPad) * bps, 0);
        if (bd > 8)
            fill_plane(r, rp, (uint16_t *)f->planes[pl].data(), (uint16_t *)f->padded[pl].data(), f->pw[pl],
                       f->ph[pl], pl ? ssx : 0, pl ? ssy : 0, blocks, bd, sigma);
        else
            fill_plane(r, rp, f->planes[pl].data(), f->padded[pl].data(), f->pw[pl], f->ph[pl], pl ? ssx : 0,
                       pl ? ssy : 0, blocks, bd, sigma);
    }
    if (cfg->monochrome) make_monochrome(f);
    return f;
}

void *r2s_plane_padded(R2SFrame *f, int pl, ptrdiff_t *stride) {
    if (stride) *stride = (ptrdiff_t)(f->pw[pl] + kPad) * (f->cfg.bitdepth > 8 ? 2 : 1);
    return f->padded[pl].data();
}

void r2s_mask_edges_part(uint16_t (*masks)[64][5][4], int by4, int bx4, int w4, int h4, int part, int tw4, int th4,
                         int hlim, int vlim, uint8_t *a, uint8_t *l) {
    edges_part(masks[0], masks[1], by4, bx4, w4, h4, part, tw4, th4, hlim, vlim, a, l);
}

void r2s_mask_subpu_edges(uint16_t (*masks)[64][5][4], int by4, int bx4, int w4, int h4, int twl4c, int thl4c,
                          int hsz, int vsz, int ds_mask) {
    subpu_edges(masks[0], masks[1], by4, bx4, w4, h4, twl4c, thl4c, hsz, vsz, ds_mask);
}

void r2s_frame_destroy(R2SFrame *f) { delete f; }
Rav2dFrameParams *r2s_params(R2SFrame *f) { return &f->params; }
void *r2s_plane(R2SFrame *f, int pl) { return f->planes[pl].data(); }
int32_t r2s_plane_width(const R2SFrame *f, int pl) { return f->pw[pl]; }
int32_t r2s_plane_height(const R2SFrame *f, int pl) { return f->ph[pl]; }
size_t r2s_plane_bytes(const R2SFrame *f, int pl) { return f->planes[pl].size(); }
Rav2dAv2Filter *r2s_mask(R2SFrame *f, size_t *n) {
    if (n) *n = f->mask.size();
    return f->mask.data();
}
uint8_t *r2s_segmap(R2SFrame *f, int which, ptrdiff_t *stride) {
    if (which < 0 || which > 1 || f->seg[which].empty()) {
        if (stride) *stride = 0;
        return nullptr;
    }
    if (stride) *stride = f->seg_stride[which];
    return f->seg[which].data();
}

Rav2dAv2Restoration *r2s_lr(R2SFrame *f, size_t *n) {
    if (n) *n = f->lr.size();
    return f->lr.data();
}

}  // extern "C"
