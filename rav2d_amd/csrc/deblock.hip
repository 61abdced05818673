// AV2 deblocking on gfx950: vertical-edge pass, then horizontal-edge pass.
//
// Arithmetic follows crates/rav2d/src/deblock.rs: filter_choice_bd (:81-186),
// deblock_bd (:217-283), the per-column/row walkers deblock_{h,v}_sb64{y,uv}_bd
// (:303-551) and the threshold set-up setup_thr_cols/rows + edge_thr
// (:1466-1567) driven by deblock64_cols/rows (:1623-2140).
//
// Layout: one workgroup filters a 64x32 (V) or 64x32 (H) tile of one plane.
// The tile plus a 16-sample halo across the edge direction is staged in LDS
// as 16-bit samples, every edge whose footprint reaches the tile is filtered
// in LDS by one lane per (edge, 4-line segment), and the tile interior is
// written to the output plane.  Edges of one pass are filtered concurrently:
// an edge of level L (transform >= 4<<L on both sides) reads at most 8
// samples on each side and its writes never reach a sample another edge of
// the same pass reads, so the result equals the reference's sequential walk.
#include "deblock_core.h"
#include "internal.h"

namespace r2g {

// deblock_bd on an LDS segment: `p` = first sample on the positive side of
// line 0, SA steps along the edge, SB across it.  The segment's four lines of
// 16 samples (-8..7 across the edge; the filters read and write at most 8 per
// side) are read into registers before any is written: the lines of a
// segment, and the segments of one pass, touch disjoint samples.  `p` is
// 4-sample aligned, so with SA = 1 (horizontal edges) the four lines' samples
// at one offset are one 8-byte word, and with SB = 1 (vertical edges) a line
// is four 8-byte words.
template <int SA, int SB>
__device__ __forceinline__ void filter_edge(uint16_t *p, int bdmax, uint32_t q_thr, uint32_t side_thr, int mwp,
                                            int mwn, bool pos_ll, bool neg_ll) {
    static_assert(SA == 1 || SB == 1, "one of the steps is a sample");
    int L[4][16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if constexpr (SA == 1) {
            const uint2 w = *(const uint2 *)&p[(k - 8) * SB];
            L[0][k] = (int)(w.x & 0xffff);
            L[1][k] = (int)(w.x >> 16);
            L[2][k] = (int)(w.y & 0xffff);
            L[3][k] = (int)(w.y >> 16);
        } else if ((k & 3) == 0) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint2 w = *(const uint2 *)&p[i * SA + k - 8];
                L[i][k] = (int)(w.x & 0xffff);
                L[i][k + 1] = (int)(w.x >> 16);
                L[i][k + 2] = (int)(w.y & 0xffff);
                L[i][k + 3] = (int)(w.y >> 16);
            }
        }
    }
    const int width = filter_choice(L[0], L[3], mwn, mwp, q_thr, side_thr);
    const int wn = dmin(width, mwn), wp = width;
    if (wp < 1) return;
    // q_thr < 2^16 and the factors below fit 24 bits (|delta| < 2^17, weights
    // <= 85, widths <= 8): 24-bit multiplies
    const int clampv = (int)__umul24(q_thr, (uint32_t)byte_at(kQThreshMults, width - 1));
    const int mneg = byte_at(kWMult, wn - 1), mpos = byte_at(kWMult, wp - 1);
    int delta[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
        delta[i] = dmax(dmin(4 * (3 * (L[i][8] - L[i][7]) - (L[i][9] - L[i][6])), clampv), -clampv);
    // tap j's correction (delta * w * (width - j) + 1024) >> 11 from a running
    // sum that steps down by delta * w per tap; the width test is shared by
    // the four lines
    auto put = [&](int o, const int (&v)[4]) {
        if constexpr (SA == 1) {
            *(uint2 *)&p[o * SB] = make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16), (uint32_t)v[2] | ((uint32_t)v[3] << 16));
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) p[i * SA + o] = (uint16_t)v[i];
        }
    };
    if (!neg_ll) {
        int dn[4], acc[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            dn[i] = __mul24(delta[i], mneg);
            acc[i] = __mul24(dn[i], wn) + 1024;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (j < wn) {
                int v[4];
#pragma unroll
                for (int i = 0; i < 4; i++) v[i] = dmax(dmin(L[i][7 - j] + (acc[i] >> 11), bdmax), 0);
                put(-j - 1, v);
            }
#pragma unroll
            for (int i = 0; i < 4; i++) acc[i] -= dn[i];
        }
    }
    if (!pos_ll) {
        int dp[4], acc[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            dp[i] = __mul24(delta[i], mpos);
            acc[i] = __mul24(dp[i], wp) + 1024;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (j < wp) {
                int v[4];
#pragma unroll
                for (int i = 0; i < 4; i++) v[i] = dmax(dmin(L[i][8 + j] - (acc[i] >> 11), bdmax), 0);
                put(j, v);
            }
#pragma unroll
            for (int i = 0; i < 4; i++) acc[i] -= dp[i];
        }
    }
}

__device__ __forceinline__ int edge_thr(int cur, int prev) {
    return (cur && prev) ? (cur + prev + 1) >> 1 : (cur | prev);
}

__device__ __forceinline__ const Rav2dPackedFilter &sb_at(const Rav2dPackedFilter *m, int sb256w, int lx4,
                                                          int ly4) {
    return m[(ly4 >> 6) * sb256w + (lx4 >> 6)];
}

// qidx of the 64x64 luma block holding luma 4px unit (lx4, ly4).
__device__ __forceinline__ int qidx_at(const Rav2dPackedFilter *m, int sb256w, int lx4, int ly4) {
    const Rav2dPackedFilter &f = sb_at(m, sb256w, lx4, ly4);
    return dmin((int)f.qidx[((ly4 >> 4) & 3) * 4 + ((lx4 >> 4) & 3)], 511);
}

struct EdgeDesc {
    bool on;
    int mwp, mwn;
    int subpu, qc, qp;  // threshold shift and the qidx of this block and of its neighbour
    bool pos_ll, neg_ll;
};

// Edge of the given pass (vert = vertical edges, cols pass) at the plane's 4px
// unit (x4, y4); pl selects luma or chroma indexing.  Every side-info load is
// issued unconditionally (one round of loads); the thresholds come later from
// the LUT rows staged in LDS.
// kAny: a frame of any chroma subsampling (4:2:2 / 4:4:4 masks in a.uvm); the
// 4:2:0 / I400 instantiation has the subsampling as constants.
template <bool kAny>
__device__ __forceinline__ EdgeDesc edge_desc(const DeblockArgs &a, int pl, bool vert, int x4, int y4) {
    if constexpr (!kAny) {
        // 4:2:0 / I400: chroma units are luma units >> 1; masks in the packed record
        EdgeDesc e;
        const int sbw = a.g.sb256w;
        const bool ch = pl != 0;
        // luma 4px units of this sample (4:2:0 chroma)
        const int lx4 = ch ? x4 * 2 : x4, ly4 = ch ? y4 * 2 : y4;
        const Rav2dPackedFilter &f = sb_at(a.mask, sbw, lx4, ly4);
        const int along = vert ? x4 : y4;   // index of the edge line within the SB256
        const int across = vert ? y4 : x4;  // position along the edge
        const int line = ch ? (along & 31) : (along & 63);
        const int sidx = ch ? ((across >> 4) & 1) : ((across >> 4) & 3);
        const uint32_t bit = 1u << (across & 15);
        // level words 0..4 of the edge line (chroma lines hold two words per level)
        const uint16_t *lw = ch ? &f.filter_uv[vert ? 0 : 1][line][0][sidx] : &f.filter_y[vert ? 0 : 1][line][0][sidx];
        const int ls = ch ? 2 : 4;
        const uint32_t m0 = lw[0], m1 = lw[ls], m2 = lw[2 * ls];
        const uint32_t m3 = ch ? 0u : (uint32_t)lw[3 * ls], m4 = lw[4 * ls];
        const int nx4 = vert ? x4 - 1 : x4, ny4 = vert ? y4 : y4 - 1;
        const int lnx4 = ch ? nx4 * 2 : nx4, lny4 = ch ? ny4 * 2 : ny4;
        e.qc = qidx_at(a.mask, sbw, lx4, ly4);
        e.qp = qidx_at(a.mask, sbw, lnx4, lny4);
        if (a.seg[0]) {
            // per-segment qidx of this unit and of its neighbour (setup_thr_cols/rows
            // read the segment map at both, deblock.rs:1477-1567)
            const uint8_t *sm = a.seg[ch ? 1 : 0];
            const int st = a.seg_stride[ch ? 1 : 0];
            e.qc = iclip(e.qc + a.seg_dq[sm[y4 * st + x4] & 7], 0, a.qmax);
            e.qp = iclip(e.qp + a.seg_dq[sm[ny4 * st + nx4] & 7], 0, a.qmax);
        }
        const Rav2dPackedFilter &fn = sb_at(a.mask, sbw, lnx4, lny4);
        if (ch) {
            e.pos_ll = (f.lossless_mask_uv[y4 & 31][(x4 >> 3) & 3] >> (x4 & 7)) & 1;
            e.neg_ll = (fn.lossless_mask_uv[ny4 & 31][(nx4 >> 3) & 3] >> (nx4 & 7)) & 1;
        } else {
            e.pos_ll = (f.lossless_mask_y[y4 & 63][(x4 >> 4) & 3] >> (x4 & 15)) & 1;
            e.neg_ll = (fn.lossless_mask_y[ny4 & 63][(nx4 >> 4) & 3] >> (nx4 & 15)) & 1;
        }
        e.on = ((m0 | m1 | m2 | m3) & bit) != 0;
        const int idx = (m3 & bit) ? 3 : (m2 & bit) ? 2 : ((m1 & bit) ? 1 : 0);
        e.mwp = byte_at(ch ? kMaxWidthUV : kMaxWidthY, idx);
        e.subpu = (m4 & bit) ? 3 : 0;
        // the rows pass clamps the negative side at the top of every 64-luma-row band
        const bool band_top = !vert && ((ch ? (y4 & 7) : (y4 & 15)) == 0);
        e.mwn = band_top ? dmin(ch ? 2 : 6, e.mwp) : e.mwp;
        return e;
    } else {
        EdgeDesc e;
        const int sbw = a.g.sb256w;
        const bool ch = pl != 0;
        // chroma subsampling (luma: none); the packed record holds 4:2:0 chroma
        // masks, 4:2:2 / 4:4:4 frames carry theirs at full size in a.uvm
        const int sx = ch ? (kAny ? a.g.ss_hor : 1) : 0, sy = ch ? (kAny ? a.g.ss_ver : 1) : 0;
        // luma 4px units of this sample
        const int lx4 = x4 << sx, ly4 = y4 << sy;
        const int sbi = (ly4 >> 6) * sbw + (lx4 >> 6);
        const Rav2dPackedFilter &f = a.mask[sbi];
        const ChromaMasks *um = (kAny && ch) ? a.uvm : nullptr;
        const int along = vert ? x4 : y4;   // index of the edge line within the SB256
        const int across = vert ? y4 : x4;  // position along the edge
        const int s_al = vert ? sx : sy, s_ac = vert ? sy : sx;
        const int line = along & ((64 >> s_al) - 1);
        const int sidx = (across >> 4) & ((4 >> s_ac) - 1);
        const uint32_t bit = 1u << (across & 15);
        // level words 0..4 of the edge line (4:2:0 chroma lines hold two words per level)
        const uint16_t *lw = um ? &um[sbi].filter_uv[vert ? 0 : 1][line][0][sidx]
                                : (ch ? &f.filter_uv[vert ? 0 : 1][line][0][sidx] : &f.filter_y[vert ? 0 : 1][line][0][sidx]);
        const int ls = (ch && !um) ? 2 : 4;
        const uint32_t m0 = lw[0], m1 = lw[ls], m2 = lw[2 * ls];
        const uint32_t m3 = ch ? 0u : (uint32_t)lw[3 * ls], m4 = lw[4 * ls];
        const int nx4 = vert ? x4 - 1 : x4, ny4 = vert ? y4 : y4 - 1;
        const int lnx4 = nx4 << sx, lny4 = ny4 << sy;
        const int nsbi = (lny4 >> 6) * sbw + (lnx4 >> 6);
        e.qc = qidx_at(a.mask, sbw, lx4, ly4);
        e.qp = qidx_at(a.mask, sbw, lnx4, lny4);
        if (a.seg[0]) {
            // per-segment qidx of this unit and of its neighbour (setup_thr_cols/rows
            // read the segment map at both, deblock.rs:1477-1567)
            const uint8_t *sm = a.seg[ch ? 1 : 0];
            const int st = a.seg_stride[ch ? 1 : 0];
            e.qc = iclip(e.qc + a.seg_dq[sm[y4 * st + x4] & 7], 0, a.qmax);
            e.qp = iclip(e.qp + a.seg_dq[sm[ny4 * st + nx4] & 7], 0, a.qmax);
        }
        const Rav2dPackedFilter &fn = a.mask[nsbi];
        if (ch) {
            // chroma lossless rows: one word per 64 luma columns, 16 >> sx units each
            const int rows = (64 >> sy) - 1, wsh = 4 - sx, bm = (16 >> sx) - 1;
            const uint16_t(*ll)[4] = um ? um[sbi].lossless_mask_uv : f.lossless_mask_uv;
            const uint16_t(*nll)[4] = um ? um[nsbi].lossless_mask_uv : fn.lossless_mask_uv;
            e.pos_ll = (ll[y4 & rows][(x4 >> wsh) & 3] >> (x4 & bm)) & 1;
            e.neg_ll = (nll[ny4 & rows][(nx4 >> wsh) & 3] >> (nx4 & bm)) & 1;
        } else {
            e.pos_ll = (f.lossless_mask_y[y4 & 63][(x4 >> 4) & 3] >> (x4 & 15)) & 1;
            e.neg_ll = (fn.lossless_mask_y[ny4 & 63][(nx4 >> 4) & 3] >> (nx4 & 15)) & 1;
        }
        e.on = ((m0 | m1 | m2 | m3) & bit) != 0;
        const int idx = (m3 & bit) ? 3 : (m2 & bit) ? 2 : ((m1 & bit) ? 1 : 0);
        e.mwp = byte_at(ch ? kMaxWidthUV : kMaxWidthY, idx);
        e.subpu = (m4 & bit) ? 3 : 0;
        // the rows pass clamps the negative side at the top of every 64-luma-row band
        const bool band_top = !vert && ((y4 & ((16 >> sy) - 1)) == 0);
        e.mwn = band_top ? dmin(ch ? 2 : 6, e.mwp) : e.mwp;
        return e;
    }
}

constexpr int kTW = 64, kTH = 32, kHalo = 16, kThreads = 256;

// Both passes: each lane issues its tile loads, its share of the threshold LUT
// rows and its edge's side-info loads in one round, then stores the tile and
// the LUT rows to LDS; thresholds are looked up there.  Every (edge line, 4-line segment) of a
// tile is one lane (152 for V, 176 for H).
template <typename P, bool kAny>
__device__ __forceinline__ void deblock_v_body(const DeblockArgs &a, int tile_id, uint16_t (*tile)[kTW + 2 * kHalo],
                                               uint16_t (*lut)[512]) {
    int pl, tx, ty;
    decode_tile(a.map, tile_id, pl, tx, ty);
    const DevPlane &src = a.src[pl];
    const DevPlane &dst = a.dst[pl];
    const int x0 = tx * kTW, y0 = ty * kTH;
    const int W = src.w, H = src.h;
    // the picture padding right of and below the plane is read too (edges of
    // transforms that overhang the frame look into it) and the rows below are
    // carried into the output for the horizontal-edge pass
    const int WP = W + a.pad_cols, HP = H + a.pad_rows;
    const int t = threadIdx.x;
    constexpr int LW = kTW + 2 * kHalo;
    // 4-sample groups (x0 - kHalo, plane widths and padding are multiples of
    // 4, so a group is wholly inside or outside), loaded from clamped, always
    // valid addresses, then selected
    constexpr int QW = LW / 4, PV = kTH * QW / kThreads;
    uint2 v[PV];
#pragma unroll
    for (int j = 0; j < PV; j++) {
        const int i = t + j * kThreads;
        const int r = i / QW, gx = x0 - kHalo + (i - r * QW) * 4, gy = y0 + r;
        const uint2 l = load4(row_ptr<const P>(src, dmin(gy, HP - 1)) + iclip(gx, 0, WP - 4));
        v[j] = (gy < HP && gx >= 0 && gx < WP) ? l : make_uint2(0, 0);
    }
    constexpr int NE = kTW / 4 + 3;  // edge columns x0-4 .. x0+TW+4
    static_assert(NE * (kTH / 4) <= kThreads, "one lane per edge segment");
    // the threshold LUT rows of this plane's kind (2 x 512 entries)
    const int kind = pl ? 1 + pl : 0;
    const uint2 lutv = ((const uint2 *)(t < 128 ? a.lut->q[kind] : a.lut->side[kind]))[t & 127];
    EdgeDesc e;
    e.on = false;
    const int rs = t / NE, ec = t - rs * NE;
    const int ex = x0 - 4 + ec * 4, ey = y0 + rs * 4;
    // no filtering at the frame's left edge
    if (a.plane_on[pl] && t < NE * (kTH / 4) && ex > 0 && ex < W && ey < H) e = edge_desc<kAny>(a, pl, true, ex >> 2, ey >> 2);
#pragma unroll
    for (int j = 0; j < PV; j++) ((uint2 *)&tile[0][0])[t + j * kThreads] = v[j];
    ((uint2 *)&lut[0][0])[t] = lutv;
    __syncthreads();
    if (e.on) {
        const uint32_t q = (uint32_t)(edge_thr(lut[0][e.qc], lut[0][e.qp]) >> e.subpu);
        const uint32_t side = (uint32_t)(edge_thr(lut[1][e.qc], lut[1][e.qp]) >> e.subpu);
        filter_edge<LW, 1>(&tile[rs * 4][ex - x0 + kHalo], a.g.bdmax, q, side, e.mwp, e.mwn, e.pos_ll, e.neg_ll);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kTH * kTW / 4 / kThreads; j++) {
        const int i = t + j * kThreads;
        const int r = i / (kTW / 4), c = (i - r * (kTW / 4)) * 4;
        const int gx = x0 + c, gy = y0 + r;
        if (gy < HP && gx < W) store4(row_ptr<P>(dst, gy) + gx, *(const uint2 *)&tile[r][c + kHalo]);
    }
}

template <typename P, bool kAny>
__device__ __forceinline__ void deblock_h_body(const DeblockArgs &a, int tile_id, uint16_t (*tile)[kTW],
                                               uint16_t (*lut)[512]) {
    int pl, tx, ty;
    decode_tile(a.map, tile_id, pl, tx, ty);
    const DevPlane &src = a.src[pl];
    const DevPlane &dst = a.dst[pl];
    const int x0 = tx * kTW, y0 = ty * kTH;
    const int W = src.w, H = src.h;
    const int HP = H + a.pad_rows;  // the padding rows below the plane are read too
    const int t = threadIdx.x;
    constexpr int LH = kTH + 2 * kHalo;
    constexpr int QW = kTW / 4, PV = LH * QW / kThreads;
    uint2 v[PV];
#pragma unroll
    for (int j = 0; j < PV; j++) {
        const int i = t + j * kThreads;
        const int r = i / QW, gx = x0 + (i - r * QW) * 4, gy = y0 - kHalo + r;
        const uint2 l = load4(row_ptr<const P>(src, iclip(gy, 0, HP - 1)) + dmin(gx, W - 4));
        v[j] = (gy >= 0 && gy < HP && gx < W) ? l : make_uint2(0, 0);
    }
    constexpr int NE = kTH / 4 + 3;  // edge rows y0-4 .. y0+TH+4
    static_assert(NE * (kTW / 4) <= kThreads, "one lane per edge segment");
    const int kind = pl ? 1 + pl : 1;
    const uint2 lutv = ((const uint2 *)(t < 128 ? a.lut->q[kind] : a.lut->side[kind]))[t & 127];
    EdgeDesc e;
    e.on = false;
    const int er = t / (kTW / 4), cs = t - er * (kTW / 4);
    const int ey = y0 - 4 + er * 4, ex = x0 + cs * 4;
    // no filtering at the frame's top edge
    if (a.plane_on[pl] && t < NE * (kTW / 4) && ey > 0 && ey < H && ex < W) e = edge_desc<kAny>(a, pl, false, ex >> 2, ey >> 2);
#pragma unroll
    for (int j = 0; j < PV; j++) ((uint2 *)&tile[0][0])[t + j * kThreads] = v[j];
    ((uint2 *)&lut[0][0])[t] = lutv;
    __syncthreads();
    if (e.on) {
        const uint32_t q = (uint32_t)(edge_thr(lut[0][e.qc], lut[0][e.qp]) >> e.subpu);
        const uint32_t side = (uint32_t)(edge_thr(lut[1][e.qc], lut[1][e.qp]) >> e.subpu);
        filter_edge<1, kTW>(&tile[ey - y0 + kHalo][cs * 4], a.g.bdmax, q, side, e.mwp, e.mwn, e.pos_ll, e.neg_ll);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kTH * kTW / 4 / kThreads; j++) {
        const int i = t + j * kThreads;
        const int r = i / (kTW / 4), c = (i - r * (kTW / 4)) * 4;
        const int gx = x0 + c, gy = y0 + r;
        if (gy < H && gx < W) store4(row_ptr<P>(dst, gy) + gx, *(const uint2 *)&tile[r + kHalo][c]);
    }
}

// Both passes on one tile of TH output rows.  The horizontal-edge filters
// whose footprint reaches rows [y0, y0 + TH) sit at rows y0 - 4 .. y0 + TH + 4
// and read 8 rows on each side, so the vertical-edge pass's output is needed
// on rows [y0 - 12, y0 + TH + 12): those rows of F0 are staged (columns with
// the vertical pass's 16-sample halo), their vertical edges filtered in LDS
// exactly as the vertical-edge pass filters them (halo rows included), then
// the horizontal edges, and the tile interior is written.  Rows outside the
// plane and its padding stage as zero, which is what the separate
// horizontal-edge pass reads there.
constexpr int kVHHalo = 12;

template <int TH> struct VHTile {
    static constexpr int LW = kTW + 2 * kHalo, R = TH + 2 * kVHHalo;
    uint16_t px[R][LW];
    uint16_t lut[4][512];  // q / side thresholds of the vertical-edge kind, then of the horizontal-edge kind
};

template <typename P, int TH, bool kAny>
__device__ __forceinline__ void deblock_vh_body(const DeblockVHArgs &f, int tile_id, VHTile<TH> &sm) {
    const DeblockArgs &a = f.a;
    constexpr int LW = VHTile<TH>::LW, R = VHTile<TH>::R;
    int pl, tx, ty;
    decode_tile(f.map, tile_id, pl, tx, ty);
    const DevPlane &src = a.src[pl];
    const DevPlane &dst = f.dst[pl];
    const int x0 = tx * kTW, y0 = ty * TH, ys = y0 - kVHHalo;  // ys: plane row of staged row 0
    const int W = src.w, H = src.h;
    const int WP = W + a.pad_cols, HP = H + a.pad_rows;
    const int t = threadIdx.x;
    constexpr int QW = LW / 4, NQ = R * QW, PV = (NQ + kThreads - 1) / kThreads;
    uint2 v[PV];
#pragma unroll
    for (int j = 0; j < PV; j++) {
        const int i = dmin(t + j * kThreads, NQ - 1);
        const int r = i / QW, gx = x0 - kHalo + (i - r * QW) * 4, gy = ys + r;
        const uint2 l = load4(row_ptr<const P>(src, iclip(gy, 0, HP - 1)) + iclip(gx, 0, WP - 4));
        v[j] = (gy >= 0 && gy < HP && gx >= 0 && gx < WP) ? l : make_uint2(0, 0);
    }
    // the four LUT rows, 16 bytes per lane
    const int kv = pl ? 1 + pl : 0, kh = pl ? 1 + pl : 1, w4 = t >> 6;
    const uint16_t *lrow = w4 == 0 ? a.lut->q[kv] : (w4 == 1 ? a.lut->side[kv] : (w4 == 2 ? a.lut->q[kh] : a.lut->side[kh]));
    const uint4 lutv = ((const uint4 *)lrow)[t & 63];
    // vertical edges: columns x0 - 4 .. x0 + TW + 4, 4-row segments of every staged row
    constexpr int NEV = kTW / 4 + 3, NSV = NEV * (R / 4), NV = (NSV + kThreads - 1) / kThreads;
    EdgeDesc ev[NV];
#pragma unroll
    for (int s = 0; s < NV; s++) {
        ev[s].on = false;
        const int idx = t + s * kThreads;
        const int rs = idx / NEV, ex = x0 - 4 + (idx - rs * NEV) * 4, ey = ys + rs * 4;
        if (f.on_v[pl] && idx < NSV && ex > 0 && ex < W && ey >= 0 && ey < H) ev[s] = edge_desc<kAny>(a, pl, true, ex >> 2, ey >> 2);
    }
#pragma unroll
    for (int j = 0; j < PV; j++)
        if (t + j * kThreads < NQ) ((uint2 *)&sm.px[0][0])[t + j * kThreads] = v[j];
    ((uint4 *)&sm.lut[0][0])[t] = lutv;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NV; s++) {
        const EdgeDesc &e = ev[s];
        if (e.on) {
            const int idx = t + s * kThreads;
            const int rs = idx / NEV, ec = idx - rs * NEV;
            const uint32_t q = (uint32_t)(edge_thr(sm.lut[0][e.qc], sm.lut[0][e.qp]) >> e.subpu);
            const uint32_t side = (uint32_t)(edge_thr(sm.lut[1][e.qc], sm.lut[1][e.qp]) >> e.subpu);
            filter_edge<LW, 1>(&sm.px[rs * 4][ec * 4 - 4 + kHalo], a.g.bdmax, q, side, e.mwp, e.mwn, e.pos_ll, e.neg_ll);
        }
    }
    __syncthreads();
    // horizontal edges: rows y0 - 4 .. y0 + TH + 4, 4-column segments of the tile
    constexpr int NEH = TH / 4 + 3, NSH = NEH * (kTW / 4), NH = (NSH + kThreads - 1) / kThreads;
#pragma unroll
    for (int s = 0; s < NH; s++) {
        const int idx = t + s * kThreads;
        const int er = idx / (kTW / 4), cs = idx - er * (kTW / 4);
        const int ey = y0 - 4 + er * 4, ex = x0 + cs * 4;
        if (f.on_h[pl] && idx < NSH && ey > 0 && ey < H && ex < W) {
            const EdgeDesc e = edge_desc<kAny>(a, pl, false, ex >> 2, ey >> 2);
            if (e.on) {
                const uint32_t q = (uint32_t)(edge_thr(sm.lut[2][e.qc], sm.lut[2][e.qp]) >> e.subpu);
                const uint32_t side = (uint32_t)(edge_thr(sm.lut[3][e.qc], sm.lut[3][e.qp]) >> e.subpu);
                filter_edge<1, LW>(&sm.px[ey - ys][cs * 4 + kHalo], a.g.bdmax, q, side, e.mwp, e.mwn, e.pos_ll,
                                   e.neg_ll);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < TH * kTW / 4 / kThreads; j++) {
        const int i = t + j * kThreads;
        const int r = i / (kTW / 4), c = (i - r * (kTW / 4)) * 4;
        const int gx = x0 + c, gy = y0 + r;
        if (gy < H && gx < W) store4(row_ptr<P>(dst, gy) + gx, *(const uint2 *)&sm.px[r + kVHHalo][c + kHalo]);
    }
}

// Kernels: kAny instantiations for frames of any chroma subsampling (the
// launchers pick them for 4:2:2 / 4:4:4 frames; a body holding both forms
// would cost the 4:2:0 form registers and occupancy).
template <typename P, int TH, bool kAny>
__global__ __launch_bounds__(kThreads) void deblock_vh_kernel(DeblockVHArgs a) {
    __shared__ __align__(16) VHTile<TH> sm;
    deblock_vh_body<P, TH, kAny>(a, blockIdx.x, sm);
}

template <typename P, int TH, bool kLoop, bool kAny>
__global__ __launch_bounds__(kThreads) void deblock_vh_batch(const DeblockVHArgs *__restrict__ args, BatchSched s) {
    __shared__ __align__(16) VHTile<TH> sm;
    for_each_item<kLoop>(s, [&](int f, int t) {
        deblock_vh_body<P, TH, kAny>(args[f], t, sm);
        __syncthreads();
    });
}

template <typename P, bool kAny> __global__ __launch_bounds__(kThreads) void deblock_v_kernel(DeblockArgs a) {
    __shared__ __align__(16) uint16_t tile[kTH][kTW + 2 * kHalo];
    __shared__ __align__(16) uint16_t lut[2][512];
    deblock_v_body<P, kAny>(a, blockIdx.x, tile, lut);
}

template <typename P, bool kAny> __global__ __launch_bounds__(kThreads) void deblock_h_kernel(DeblockArgs a) {
    __shared__ __align__(16) uint16_t tile[kTH + 2 * kHalo][kTW];
    __shared__ __align__(16) uint16_t lut[2][512];
    deblock_h_body<P, kAny>(a, blockIdx.x, tile, lut);
}

// Batched launches: every tile of every frame of one bit depth.
template <typename P, bool kLoop, bool kAny>
__global__ __launch_bounds__(kThreads) void deblock_v_batch(const DeblockArgs *__restrict__ args, BatchSched s) {
    __shared__ __align__(16) uint16_t tile[kTH][kTW + 2 * kHalo];
    __shared__ __align__(16) uint16_t lut[2][512];
    for_each_item<kLoop>(s, [&](int f, int t) {
        deblock_v_body<P, kAny>(args[f], t, tile, lut);
        __syncthreads();
    });
}

template <typename P, bool kLoop, bool kAny>
__global__ __launch_bounds__(kThreads) void deblock_h_batch(const DeblockArgs *__restrict__ args, BatchSched s) {
    __shared__ __align__(16) uint16_t tile[kTH + 2 * kHalo][kTW];
    __shared__ __align__(16) uint16_t lut[2][512];
    for_each_item<kLoop>(s, [&](int f, int t) {
        deblock_h_body<P, kAny>(args[f], t, tile, lut);
        __syncthreads();
    });
}

// Mixed launches: the tiles of every bit-depth group's frames in one grid;
// group 0 (8-bit) runs the 8-bit body, groups 1 and 2 the 16-bit one.
template <bool kVert, bool kLoop, bool kAny>
__global__ __launch_bounds__(kThreads) void deblock_mix(DevArgs3<DeblockArgs> args, MixSched s) {
    __shared__ __align__(16) uint16_t tile[kVert ? kTH : kTH + 2 * kHalo][kVert ? kTW + 2 * kHalo : kTW];
    __shared__ __align__(16) uint16_t lut[2][512];
    for_each_mixed_item<kLoop>(s, [&](int g, int f, int t) {
        if constexpr (kVert) {
            if (g == 0)
                deblock_v_body<uint8_t, kAny>(args.p[0][f], t, tile, lut);
            else
                deblock_v_body<uint16_t, kAny>(args.p[g][f], t, tile, lut);
        } else {
            if (g == 0)
                deblock_h_body<uint8_t, kAny>(args.p[0][f], t, tile, lut);
            else
                deblock_h_body<uint16_t, kAny>(args.p[g][f], t, tile, lut);
        }
        __syncthreads();
    });
}

template <int TH, bool kLoop, bool kAny>
__global__ __launch_bounds__(kThreads) void deblock_vh_mix(DevArgs3<DeblockVHArgs> args, MixSched s) {
    __shared__ __align__(16) VHTile<TH> sm;
    for_each_mixed_item<kLoop>(s, [&](int g, int f, int t) {
        if (g == 0)
            deblock_vh_body<uint8_t, TH, kAny>(args.p[0][f], t, sm);
        else
            deblock_vh_body<uint16_t, TH, kAny>(args.p[g][f], t, sm);
        __syncthreads();
    });
}

// Picks the instantiation of launch expression L for the flags given.
#define R2G_ANY(any, L)              \
    do {                             \
        if (any) {                   \
            constexpr bool kA = true;  \
            L;                       \
        } else {                     \
            constexpr bool kA = false; \
            L;                       \
        }                            \
    } while (0)

void launch_deblock_mix(bool vert, DevArgs3<DeblockArgs> args, const MixSched &s, int grid, bool any, hipStream_t st) {
    if (s.first[3] <= 0) return;
    const bool loop = grid < s.first[3];
    const dim3 G(grid), B(kThreads);
    if (vert) {
        if (loop)
            R2G_ANY(any, R2G_LAUNCH((deblock_mix<true, true, kA>), G, B, 0, st, args, s));
        else
            R2G_ANY(any, R2G_LAUNCH((deblock_mix<true, false, kA>), G, B, 0, st, args, s));
    } else {
        if (loop)
            R2G_ANY(any, R2G_LAUNCH((deblock_mix<false, true, kA>), G, B, 0, st, args, s));
        else
            R2G_ANY(any, R2G_LAUNCH((deblock_mix<false, false, kA>), G, B, 0, st, args, s));
    }
}

void launch_deblock_vh_mix(DevArgs3<DeblockVHArgs> args, const MixSched &s, int grid, int rows, bool any,
                           hipStream_t st) {
    if (s.first[3] <= 0) return;
    const bool loop = grid < s.first[3];
    const dim3 G(grid), B(kThreads);
    if (rows == 64) {
        if (loop)
            R2G_ANY(any, R2G_LAUNCH((deblock_vh_mix<64, true, kA>), G, B, 0, st, args, s));
        else
            R2G_ANY(any, R2G_LAUNCH((deblock_vh_mix<64, false, kA>), G, B, 0, st, args, s));
    } else {
        if (loop)
            R2G_ANY(any, R2G_LAUNCH((deblock_vh_mix<32, true, kA>), G, B, 0, st, args, s));
        else
            R2G_ANY(any, R2G_LAUNCH((deblock_vh_mix<32, false, kA>), G, B, 0, st, args, s));
    }
}

void launch_deblock_batch(bool vert, const DeblockArgs *d_args, const BatchSched &s, int grid, int bitdepth, bool any,
                          hipStream_t st) {
    if (s.total <= 0) return;
    const bool loop = grid < s.total;
#define R2G_DB_LAUNCH(K, P, L) R2G_ANY(any, R2G_LAUNCH((K<P, L, kA>), dim3(grid), dim3(kThreads), 0, st, d_args, s))
#define R2G_DB_PICK(K)                                     \
    do {                                                   \
        if (bitdepth > 8) {                                \
            if (loop)                                      \
                R2G_DB_LAUNCH(K, uint16_t, true);          \
            else                                           \
                R2G_DB_LAUNCH(K, uint16_t, false);         \
        } else {                                           \
            if (loop)                                      \
                R2G_DB_LAUNCH(K, uint8_t, true);           \
            else                                           \
                R2G_DB_LAUNCH(K, uint8_t, false);          \
        }                                                  \
    } while (0)
    if (vert)
        R2G_DB_PICK(deblock_v_batch);
    else
        R2G_DB_PICK(deblock_h_batch);
#undef R2G_DB_PICK
#undef R2G_DB_LAUNCH
}

TileMap deblock_vh_tile_map(const DeblockArgs &a, int rows) {
    TileMap m;
    m.first[0] = 0;
    for (int p = 0; p < 3; p++) {
        m.tiles_x[p] = (a.src[p].w + kTW - 1) / kTW;
        m.tiles_y[p] = (a.src[p].h + rows - 1) / rows;
        m.first[p + 1] = m.first[p] + m.tiles_x[p] * m.tiles_y[p];
    }
    return m;
}

void launch_deblock_vh(const DeblockVHArgs &a, int bitdepth, int rows, hipStream_t s) {
    if (a.map.first[3] == 0) return;
    const dim3 grid(a.map.first[3]), block(kThreads);
    const bool any = a.a.uvm != nullptr;
    if (bitdepth > 8) {
        if (rows == 64)
            R2G_ANY(any, R2G_LAUNCH((deblock_vh_kernel<uint16_t, 64, kA>), grid, block, 0, s, a));
        else
            R2G_ANY(any, R2G_LAUNCH((deblock_vh_kernel<uint16_t, 32, kA>), grid, block, 0, s, a));
    } else {
        if (rows == 64)
            R2G_ANY(any, R2G_LAUNCH((deblock_vh_kernel<uint8_t, 64, kA>), grid, block, 0, s, a));
        else
            R2G_ANY(any, R2G_LAUNCH((deblock_vh_kernel<uint8_t, 32, kA>), grid, block, 0, s, a));
    }
}

void launch_deblock_vh_batch(const DeblockVHArgs *d_args, const BatchSched &s, int grid, int bitdepth, int rows,
                             bool any, hipStream_t st) {
    if (s.total <= 0) return;
    const bool loop = grid < s.total;
#define R2G_VH(P, TH)                                                                                        \
    do {                                                                                                     \
        if (loop)                                                                                            \
            R2G_ANY(any, R2G_LAUNCH((deblock_vh_batch<P, TH, true, kA>), dim3(grid), dim3(kThreads), 0, st,  \
                                    d_args, s));                                                             \
        else                                                                                                 \
            R2G_ANY(any, R2G_LAUNCH((deblock_vh_batch<P, TH, false, kA>), dim3(grid), dim3(kThreads), 0, st, \
                                    d_args, s));                                                             \
    } while (0)
    if (bitdepth > 8) {
        if (rows == 64)
            R2G_VH(uint16_t, 64);
        else
            R2G_VH(uint16_t, 32);
    } else {
        if (rows == 64)
            R2G_VH(uint8_t, 64);
        else
            R2G_VH(uint8_t, 32);
    }
#undef R2G_VH
}

TileMap deblock_tile_map(const DeblockArgs &a, bool vert) {
    TileMap m;
    m.first[0] = 0;
    for (int p = 0; p < 3; p++) {
        m.tiles_x[p] = (a.src[p].w + kTW - 1) / kTW;
        m.tiles_y[p] = (a.src[p].h + (vert ? a.pad_rows : 0) + kTH - 1) / kTH;
        m.first[p + 1] = m.first[p] + m.tiles_x[p] * m.tiles_y[p];
    }
    return m;
}

void launch_deblock_v(const DeblockArgs &a0, int bitdepth, hipStream_t s) {
    DeblockArgs a = a0;
    a.map = deblock_tile_map(a, true);
    if (a.map.first[3] == 0) return;
    const bool any = a.uvm != nullptr;
    if (bitdepth > 8)
        R2G_ANY(any, R2G_LAUNCH((deblock_v_kernel<uint16_t, kA>), dim3(a.map.first[3]), dim3(kThreads), 0, s, a));
    else
        R2G_ANY(any, R2G_LAUNCH((deblock_v_kernel<uint8_t, kA>), dim3(a.map.first[3]), dim3(kThreads), 0, s, a));
}

void launch_deblock_h(const DeblockArgs &a0, int bitdepth, hipStream_t s) {
    DeblockArgs a = a0;
    a.map = deblock_tile_map(a, false);
    if (a.map.first[3] == 0) return;
    const bool any = a.uvm != nullptr;
    if (bitdepth > 8)
        R2G_ANY(any, R2G_LAUNCH((deblock_h_kernel<uint16_t, kA>), dim3(a.map.first[3]), dim3(kThreads), 0, s, a));
    else
        R2G_ANY(any, R2G_LAUNCH((deblock_h_kernel<uint8_t, kA>), dim3(a.map.first[3]), dim3(kThreads), 0, s, a));
}

}  // namespace r2g
