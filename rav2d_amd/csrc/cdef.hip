// CDEF with CCSO on gfx950: one workgroup per 64x64 luma superblock (and the
// chroma of its area, in 32x32 sub-tiles), F2 -> F3.
//
// Arithmetic follows crates/rav2d/src/cdef.rs: cdef_find_dir_bd (:134-216),
// adjust_strength (:226-236), constrain (:117-120), cdef_filter_block with its
// padding (:37-128, 320-460) and the per-band/superblock/block decisions of
// cdef_brow (:702-1182); CCSO from crates/rav2d/src/ccso.rs: ccso_prep
// (:115-186), ccso_score (:16-24) and ccso_add (:215-254).
//
// The reference walks 8-row bands and keeps pre-filter copies of the rows
// above (toggled line banks) and of the two columns to the left (2x8 backups),
// so every tap it reads is a deblocked (F2) sample and every sample outside
// the frame is the INT16_MIN "no edge" marker.  Here the whole superblock
// plus a 2-sample halo of F2 is staged in LDS once, with INT16_MIN outside the
// frame, and every output sample is computed from it directly.  CCSO reads
// the same staged luma with coordinates clamped to the frame, which is what
// ccso_padding's edge-dependent clamps amount to.
#include "internal.h"
#include "tables.h"

namespace r2g {


__device__ __forceinline__ int constrain(int diff, int threshold, int shift) {
    int adiff = abs(diff);
    return apply_sign(dmin(adiff, dmax(0, threshold - (adiff >> shift))), diff);
}

__device__ __forceinline__ int adjust_strength(int strength, uint32_t var) {
    if (var == 0) return 0;
    int i = (var >> 6) ? dmin(ulog2(var >> 6), 12) : 0;
    return (strength * (4 + i) + 8) >> 4;
}

__device__ __forceinline__ uint32_t ccso_score(int diff, int q, uint32_t edge_clf) {
    if (diff > q && edge_clf == 0) return 2;
    if (diff < -q) return 0;
    return 1;
}

// Direction search of the 8x8 luma blocks, spread over the four waves of a
// workgroup: wave `role` computes, for the block of each lane, the costs of
// two orthogonal directions (d, d + 4) -- role 0: 0/4 (diagonal line sums),
// 1: 2/6 (rows, columns), 2: 1/5, 3: 3/7 (the half-slope families).  The line
// sums accumulate (px >> bdm8) as read and subtract 128 x (samples in the
// line) afterwards, which equals summing (px >> bdm8) - 128 per sample.
__device__ __forceinline__ uint32_t dir_sq(int v) {  // |v| <= 8 x 128: a 24-bit product
    return (uint32_t)((int)(int16_t)v * (int)(int16_t)v);
}

// cost of the half-slope family with 11 line sums `al` (each biased by 128 x count)
template <int N> __device__ __forceinline__ uint32_t alt_cost(const int (&al)[N], const int (&cnt)[N]) {
    constexpr uint32_t DIV[7] = {840, 420, 280, 210, 168, 140, 120};
    uint32_t c = 0;
#pragma unroll
    for (int m = 0; m < 5; m++) c += dir_sq(al[3 + m] - 128 * cnt[3 + m]);
    c *= 105;
#pragma unroll
    for (int m = 0; m < 3; m++)
        c += (dir_sq(al[m] - 128 * cnt[m]) + dir_sq(al[10 - m] - 128 * cnt[10 - m])) * DIV[2 * m + 1];
    return c;
}

template <int kRole>
__device__ __forceinline__ void dir_costs(const int16_t *img, int ls, int bdm8, uint32_t &c_lo, uint32_t &c_hi) {
    constexpr uint32_t DIV[7] = {840, 420, 280, 210, 168, 140, 120};
    int s0[15] = {0}, s1[15] = {0};
#pragma unroll
    for (int y = 0; y < 8; y++) {
        // one row: 8 samples in two 8-byte LDS reads (rows are 8-byte aligned)
        const uint2 lo = *(const uint2 *)&img[y * ls], hi = *(const uint2 *)&img[y * ls + 4];
        const uint32_t wv[4] = {lo.x, lo.y, hi.x, hi.y};
#pragma unroll
        for (int x = 0; x < 8; x++) {
            const int px = (int)((wv[x >> 1] >> (16 * (x & 1))) & 0xffff) >> bdm8;
            if constexpr (kRole == 0) {
                s0[y + x] += px;
                s1[7 + y - x] += px;
            } else if constexpr (kRole == 1) {
                s0[y] += px;
                s1[x] += px;
            } else if constexpr (kRole == 2) {
                s0[y + (x >> 1)] += px;
                s1[3 - (y >> 1) + x] += px;
            } else {
                s0[3 + y - (x >> 1)] += px;
                s1[(y >> 1) + x] += px;
            }
        }
    }
    if constexpr (kRole == 0) {
        uint32_t c0 = 0, c1 = 0;
#pragma unroll
        for (int n = 0; n < 7; n++) {
            const int cn = n + 1;  // samples on the diagonals n and 14 - n
            c0 += (dir_sq(s0[n] - 128 * cn) + dir_sq(s0[14 - n] - 128 * cn)) * DIV[n];
            c1 += (dir_sq(s1[n] - 128 * cn) + dir_sq(s1[14 - n] - 128 * cn)) * DIV[n];
        }
        c_lo = c0 + dir_sq(s0[7] - 1024) * 105;
        c_hi = c1 + dir_sq(s1[7] - 1024) * 105;
    } else if constexpr (kRole == 1) {
        uint32_t c0 = 0, c1 = 0;
#pragma unroll
        for (int n = 0; n < 8; n++) {
            c0 += dir_sq(s0[n] - 1024);
            c1 += dir_sq(s1[n] - 1024);
        }
        c_lo = c0 * 105;
        c_hi = c1 * 105;
    } else {
        // samples per line of a half-slope family: 2 x |{j in 0..3 : 0 <= n - j <= 7}|
        constexpr int cnt[11] = {2, 4, 6, 8, 8, 8, 8, 8, 6, 4, 2};
        int a0[11], a1[11];
#pragma unroll
        for (int n = 0; n < 11; n++) {
            a0[n] = s0[n];
            a1[n] = s1[n];
        }
        c_lo = alt_cost(a0, cnt);
        c_hi = alt_cost(a1, cnt);
    }
}

// Two horizontally adjacent output samples of cdef_filter_block in packed
// 16-bit lanes.  `t` points at the left sample in an LDS tile whose
// outside-frame entries hold INT16_MIN; `op`, `os0`, `os1` are the tap offsets
// (in that tile's stride) of the primary direction and of the two secondary
// directions, for taps k = 0, 1.  16 bits hold every intermediate: tap
// differences are formed with a saturating subtract, so a difference with an
// INT16_MIN neighbour stays at least 32767 in magnitude and constrains to 0
// exactly as in 32-bit arithmetic (samples below 2^12, primary strengths at
// most 15 << 4 and secondary 4 << 4 at 12 bit, shifts at most 10);
// |sum| <= 2 * 7 * 240 + 12 * 64 < 2^15.
typedef short s16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 splat2(int v) { return (s16x2){(short)v, (short)v}; }

// constrain() of N taps' differences d, in place, stage by stage across the
// taps: consecutive packed operations are then independent (a dependent pair
// of packed operations costs a wait state).
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

template <int N>
__device__ __forceinline__ void constrain_n(s16x2 (&d)[N], s16x2 thr, s16x2 sh) {
    const s16x2 z = {0, 0};
    s16x2 c[N];
#pragma unroll
    for (int k = 0; k < N; k++) c[k] = __builtin_elementwise_sub_sat(z, d[k]);
#pragma unroll
    for (int k = 0; k < N; k++) c[k] = __builtin_elementwise_max(d[k], c[k]);  // |d|
#pragma unroll
    for (int k = 0; k < N; k++) c[k] = c[k] >> sh;
    // max(0, thr - (|d| >> sh)) as one unsigned saturating subtract (both >= 0)
#pragma unroll
    for (int k = 0; k < N; k++)
        c[k] = __builtin_bit_cast(s16x2, __builtin_elementwise_sub_sat(__builtin_bit_cast(u16x2, thr),
                                                                       __builtin_bit_cast(u16x2, c[k])));
#pragma unroll
    for (int k = 0; k < N; k++) d[k] = __builtin_elementwise_max(d[k], z - c[k]);
#pragma unroll
    for (int k = 0; k < N; k++) d[k] = __builtin_elementwise_min(d[k], c[k]);
}

__device__ __forceinline__ s16x2 cdef_round(s16x2 px, s16x2 sum) {
    return px + ((sum + (sum >> (short)15) + (short)8) >> (short)4);
}

__device__ __forceinline__ s16x2 cdef_px2(const int16_t *t, const int op[2], const int os0[2], const int os1[2],
                                          int pri, int pri_tap, int pri_shift, int sec, int sec_shift) {
    // two 16-bit LDS reads per pair: odd tap offsets would make a merged 32-bit
    // read unaligned, so the reads are relaxed atomics, which are not merged.
    // They keep their source order, so every tap is read before the first use.
    auto ld = [&](int o) -> s16x2 { return (s16x2){lds_load_i16(t + o), lds_load_i16(t + o + 1)}; };
    const s16x2 ps = splat2(pri), pss = splat2(pri_shift), ss = splat2(sec), sss = splat2(sec_shift);
    // primary taps: k = 0 (weight pri_tap), k = 1 (weight (pri_tap & 3) | 2);
    // secondary taps: k = 0 (weight 2), k = 1 (weight 1).  16-bit sums wrap
    // like the reference's (|sum| < 2^15), so their order does not matter.
    const s16x2 tk0 = splat2(pri_tap), tk1 = splat2((pri_tap & 3) | 2);
    auto load_pri = [&](s16x2(&pv)[4]) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            pv[2 * k] = ld(op[k]);
            pv[2 * k + 1] = ld(-op[k]);
        }
    };
    auto load_sec = [&](s16x2(&sv)[8]) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            sv[4 * k] = ld(os0[k]);
            sv[4 * k + 1] = ld(-os0[k]);
            sv[4 * k + 2] = ld(os1[k]);
            sv[4 * k + 3] = ld(-os1[k]);
        }
    };
    auto pri_sum = [&](const s16x2(&c)[4]) { return tk0 * (c[0] + c[1]) + tk1 * (c[2] + c[3]); };
    auto sec_sum = [&](const s16x2(&c)[8]) {
        return (((c[0] + c[1]) + (c[2] + c[3])) << (short)1) + ((c[4] + c[5]) + (c[6] + c[7]));
    };
    if (pri && sec) {
        const s16x2 px = ld(0);
        s16x2 pv[4], sv[8], dp[4], dsv[8];
        load_pri(pv);
        load_sec(sv);
#pragma unroll
        for (int k = 0; k < 4; k++) dp[k] = __builtin_elementwise_sub_sat(pv[k], px);
#pragma unroll
        for (int k = 0; k < 8; k++) dsv[k] = __builtin_elementwise_sub_sat(sv[k], px);
        constrain_n(dp, ps, pss);
        constrain_n(dsv, ss, sss);
        // clamp to the range of the taps read (and the sample itself)
        s16x2 mn[6], mx[6];
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const s16x2 a = k < 2 ? pv[2 * k] : sv[2 * k - 4], b = k < 2 ? pv[2 * k + 1] : sv[2 * k - 3];
            mn[k] = __builtin_elementwise_min(a, b);
            mx[k] = __builtin_elementwise_max(a, b);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            mn[k] = __builtin_elementwise_min(mn[k], mn[k + 3]);
            mx[k] = __builtin_elementwise_max(mx[k], mx[k + 3]);
        }
        const s16x2 lo = __builtin_elementwise_min(__builtin_elementwise_min(mn[0], mn[1]), __builtin_elementwise_min(mn[2], px));
        const s16x2 hi = __builtin_elementwise_max(__builtin_elementwise_max(mx[0], mx[1]), __builtin_elementwise_max(mx[2], px));
        const s16x2 v = cdef_round(px, pri_sum(dp) + sec_sum(dsv));
        return __builtin_elementwise_min(__builtin_elementwise_max(v, lo), hi);
    }
    if (pri) {
        const s16x2 px = ld(0);
        s16x2 pv[4];
        load_pri(pv);
#pragma unroll
        for (int k = 0; k < 4; k++) pv[k] = __builtin_elementwise_sub_sat(pv[k], px);
        constrain_n(pv, ps, pss);
        return cdef_round(px, pri_sum(pv));
    }
    const s16x2 px = ld(0);
    s16x2 sv[8];
    load_sec(sv);
#pragma unroll
    for (int k = 0; k < 8; k++) sv[k] = __builtin_elementwise_sub_sat(sv[k], px);
    constrain_n(sv, ss, sss);
    return cdef_round(px, sec_sum(sv));
}

constexpr int kThreads = 256;
// Staged tiles: rows with a 2-sample halo; columns with a 4-sample halo so
// that rows are loaded 4 samples per access (the filters read 2 beyond).
constexpr int LYR = 64 + 4, LCR = 32 + 4;  // staged rows
constexpr int LY = 64 + 8, LC = 32 + 8;    // staged row strides
constexpr int CX = 4;                      // column of tile sample 0

struct CdefSmem {
    alignas(16) int16_t ly[LYR * LY];
    alignas(16) int16_t lc[2][LCR * LC];
    uint8_t b_state[64];   // bit0 luma filtered, bit1 chroma filtered
    uint8_t b_dir[64];
    int16_t b_ypri[64];    // adjusted luma primary strength
    alignas(4) int16_t off[2][12][2];    // CDEF_DIRECTIONS as offsets in the luma / chroma tile
    alignas(4) int8_t ccso_lut[3][256];  // CCSO offset per band/edge index, per plane
    uint32_t dcost[8][64]; // direction-search cost of each direction per block
    uint16_t ll[32];       // lossless 4x4 units of the superblock: 16 luma rows, 16 >> ss_ver chroma rows
};

// CCSO offset of one sample: `lt` = staged luma at the collocated position
// (lx, lyy in frame coordinates), neighbours clamped to the frame; `lut` is
// the plane's offset for each of the 256 band/edge indices, staged in LDS.
__device__ __forceinline__ int ccso_offset(const CcsoDev &c, const int8_t *lut, const int16_t *ly, int lx, int lyy,
                                           int x0, int y0, int fw, int fh) {
    const int cval = ly[(lyy - y0 + 2) * LY + lx - x0 + CX];
    int idx = (int)((uint32_t)cval >> c.shift);
    if (!c.bo_only) {
        int nx0 = iclip(lx + c.dx, 0, fw - 1) - x0 + CX, ny0 = iclip(lyy + c.dy, 0, fh - 1) - y0 + 2;
        int nx1 = iclip(lx - c.dx, 0, fw - 1) - x0 + CX, ny1 = iclip(lyy - c.dy, 0, fh - 1) - y0 + 2;
        uint32_t c0 = ccso_score(ly[ny0 * LY + nx0] - cval, c.quant_step, c.edge_clf);
        uint32_t c1 = ccso_score(ly[ny1 * LY + nx1] - cval, c.quant_step, c.edge_clf);
        idx |= (int)((c0 << 5) | (c1 << 3));
    }
    return lut[idx & 0xff];
}

// CCSO offsets of two horizontally adjacent samples whose collocated luma and
// its neighbours at +-(dx, dy) all lie inside the frame (no clamping): `cp`
// points at the collocated luma of the left sample in the staged tile, the
// right sample's is kStep further (1 luma, 2 chroma), `o` = dy * LY + dx.
// ccso_score from the sign bits of packed 16-bit differences (|d| <= 1023):
// 1 - [d < -q] + [d > q] (the last only without edge_clf), as
// 1 + ((d + q) >> 15) - ((q - d) >> 15).
template <int kStep>
__device__ __forceinline__ void ccso_pair_inner(const CcsoDev &c, const int8_t *lut, const int16_t *cp, int o,
                                                int &off0, int &off1) {
    const int c0 = cp[0], c1 = cp[kStep];
    uint32_t idx = ((uint32_t)c0 >> c.shift) | (((uint32_t)c1 >> c.shift) << 16);
    if (!c.bo_only) {
        const s16x2 cv = {(short)c0, (short)c1};
        const s16x2 na = {(short)cp[o], (short)cp[o + kStep]}, nb = {(short)cp[-o], (short)cp[kStep - o]};
        const s16x2 q = splat2(c.quant_step), one = {1, 1};
        const s16x2 da = na - cv, db = nb - cv;
        s16x2 sa = one + ((da + q) >> (short)15), sb = one + ((db + q) >> (short)15);
        if (c.edge_clf == 0) {
            sa -= (q - da) >> (short)15;
            sb -= (q - db) >> (short)15;
        }
        // (cls0 << 5) | (cls1 << 3) | band, the band in bits 0..15 of each half
        idx |= __builtin_bit_cast(uint32_t, (s16x2)((sa << (short)5) + (sb << (short)3)));
    }
    off0 = lut[idx & 0xff];
    off1 = lut[(idx >> 16) & 0xff];
}

// Staging of one superblock, split so that a workgroup can have the next
// superblock's loads in flight while it filters the current one.
constexpr int kQY = LY / 4, kQC = LC / 4;
constexpr int kNY = LYR * kQY, kNC = 2 * LCR * kQC;
constexpr int kPY = (kNY + kThreads - 1) / kThreads, kPC = (kNC + kThreads - 1) / kThreads;

struct CdefLoads {
    uint2 vy[kPY], vc[kPC];  // F2 groups of the luma / chroma tiles (INT16_MIN outside the frame)
    uint32_t lutw, offw;     // words of the CCSO offset tables and of the tap offsets
    int cdef_idx, ccso_pl;   // side info of the superblock
    uint32_t noskip_b;       // noskip row of the 8x8 block of lane t & 63 (block row (t >> 3) & 7)
    uint32_t llw;            // lanes 0..31: one row of the superblock's lossless masks
};

constexpr uint32_t kOutside = 0x80008000u;  // INT16_MIN x 2

// F2 groups of one 32x32 chroma sub-tile of both chroma planes (with its
// 2-row / 4-column halo), at chroma sample (cx0, cy0), for LDS tile lc.
template <typename P>
__device__ __forceinline__ void cdef_load_chroma(const CdefArgs &a, int cx0, int cy0, int cw, int ch, uint2 (&vc)[kPC]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int j = 0; j < kPC; j++) {
        const int i = t + j * kThreads;
        const int pl = i >= LCR * kQC;
        const int k = i - pl * LCR * kQC;
        const int r = k / kQC, gx = cx0 - CX + (k - r * kQC) * 4, gy = cy0 - 2 + r;
        uint2 v = make_uint2(kOutside, kOutside);
        if (a.has_chroma)  // monochrome frames have no chroma planes
            v = load4(row_ptr<const P>(a.src[1 + (pl & 1)], iclip(gy, 0, ch - 1)) + iclip(gx, 0, cw - 4));
        const bool in = gx >= 0 && gx < cw && gy >= 0 && gy < ch;
        vc[j].x = in ? v.x : kOutside;
        vc[j].y = in ? v.y : kOutside;
    }
}

// Every global load of a lane for superblock (sbx, sby), issued together.
// F2 is read in 4-sample groups (tile origins and plane widths are multiples
// of 4, so a group is wholly inside or outside the frame) from clamped,
// always valid addresses, then selected.
// kAny: a frame of any chroma subsampling (4:2:2 / 4:4:4 masks in a.uvm); the
// 4:2:0 / I400 instantiation has the subsampling as constants.
template <typename P, bool kAny>
__device__ __forceinline__ void cdef_load(const CdefArgs &a, int sbx, int sby, CdefLoads &L) {
    const int t = threadIdx.x;
    const int x0 = sbx * 64, y0 = sby * 64;
    const int fw = a.g.bw * 4, fh = a.g.bh * 4;
    const int ssx = kAny ? a.g.ss_hor : 1, ssy = kAny ? a.g.ss_ver : 1;
    const int cw = fw >> ssx, ch = fh >> ssy;
    const int sbi = (sby >> 2) * a.g.sb256w + (sbx >> 2);
    const Rav2dPackedFilter &m = a.mask[sbi];
    L.cdef_idx = m.cdef_idx[((sby & 3) << 2) + (sbx & 3)];
    L.ccso_pl = (m.ccso[0] != 0) | ((m.ccso[1] != 0) << 1) | ((m.ccso[2] != 0) << 2);
    L.noskip_b = a.on_skiptx ? 0xffffu : m.noskip_mask[(((y0 >> 3) + ((t >> 3) & 7)) & 31)][sbx & 3];
    // the lossless rows are read only when a segment is lossless (cdef.rs:894-911):
    // lanes 0..15 luma, 16.. the 16 >> ss_ver chroma rows of the superblock
    L.llw = 0;
    const int uvr = 16 >> ssy;
    if (a.any_ll && t < 16 + uvr) {
        const uint16_t(*uvll)[4] = (kAny && a.uvm) ? a.uvm[sbi].lossless_mask_uv : m.lossless_mask_uv;
        L.llw = t < 16 ? m.lossless_mask_y[((sby & 3) << 4) + t][sbx & 3] : uvll[(sby & 3) * uvr + t - 16][sbx & 3];
    }
#pragma unroll
    for (int j = 0; j < kPY; j++) {
        const int i = t + j * kThreads;
        const int r = i / kQY, gx = x0 - CX + (i - r * kQY) * 4, gy = y0 - 2 + r;
        const uint2 v = load4(row_ptr<const P>(a.src[0], iclip(gy, 0, fh - 1)) + iclip(gx, 0, fw - 4));
        const bool in = gx >= 0 && gx < fw && gy >= 0 && gy < fh;
        L.vy[j].x = in ? v.x : kOutside;  // per word: a vector select would go through scratch
        L.vy[j].y = in ? v.y : kOutside;
    }
    cdef_load_chroma<P>(a, x0 >> ssx, y0 >> ssy, cw, ch, L.vc);
    // the CCSO offset tables (3 x 256 B) and the tap offsets (96 B) of the argument block
    L.lutw = t < 3 * 64 ? ((const uint32_t *)a.ccso[t >> 6].lut)[t & 63] : 0u;
    L.offw = t < 24 ? ((const uint32_t *)&a.off[0][0][0])[t] : 0u;
}

// The LDS stores of a superblock's staging (after every lane is done with the
// previous superblock's LDS contents).
__device__ __forceinline__ void cdef_store(const CdefLoads &L, CdefSmem &sm) {
    const int t = threadIdx.x;
#pragma unroll
    for (int j = 0; j < kPY; j++)
        if (t + j * kThreads < kNY) ((uint2 *)sm.ly)[t + j * kThreads] = L.vy[j];
#pragma unroll
    for (int j = 0; j < kPC; j++)
        if (t + j * kThreads < kNC) ((uint2 *)&sm.lc[0][0])[t + j * kThreads] = L.vc[j];
    if (t < 3 * 64) ((uint32_t *)&sm.ccso_lut[0][0])[t] = L.lutw;
    if (t < 24) ((uint32_t *)&sm.off[0][0][0])[t] = L.offw;
    if (t < 32) sm.ll[t] = (uint16_t)L.llw;
}

// Direction search, filtering and CCSO of a superblock staged by cdef_store.
template <typename P, bool kAny>
__device__ __forceinline__ void cdef_filter(const CdefArgs &a, int sbx, int sby, const CdefLoads &L, CdefSmem &sm) {
    int16_t *ly = sm.ly;
    int16_t(*lc)[LCR * LC] = sm.lc;
    uint8_t *b_state = sm.b_state, *b_dir = sm.b_dir;
    int16_t *b_ypri = sm.b_ypri;
    const int t = threadIdx.x;
    const int x0 = sbx * 64, y0 = sby * 64;
    const int fw = a.g.bw * 4, fh = a.g.bh * 4;
    const int ssx = kAny ? a.g.ss_hor : 1, ssy = kAny ? a.g.ss_ver : 1;
    const int cw = fw >> ssx, ch = fh >> ssy;
    const int bdm8 = a.g.bitdepth - 8;
    const uint64_t y_str = *(const uint64_t *)a.y_strength, uv_str = *(const uint64_t *)a.uv_strength;
    const uint32_t noskip_b = L.noskip_b;
    const int cdef_idx_v = L.cdef_idx, ccso_pl_v = L.ccso_pl;
    // uniform side info (read after the staging loads are in flight)
    const int cdef_idx = __builtin_amdgcn_readfirstlane(cdef_idx_v);
    const int ccso_pl = __builtin_amdgcn_readfirstlane(ccso_pl_v);
    const int y_s = cdef_idx >= 0 ? (int)((y_str >> (8 * (cdef_idx & 7))) & 0xff) : 0;
    const int uv_s = cdef_idx >= 0 ? (int)((uv_str >> (8 * (cdef_idx & 7))) & 0xff) : 0;
    const bool active = a.cdef_on && cdef_idx >= 0 && (y_s != 0 || uv_s != 0);
    const int y_lvl = active ? y_s : 0;
    const int uv_lvl = active ? uv_s : 0;
    const int y_pri = (y_lvl >> 2) << bdm8;
    int y_sec = y_lvl & 3;
    y_sec += y_sec == 3;
    y_sec <<= bdm8;
    const int uv_pri = (uv_lvl >> 2) << bdm8;
    int uv_sec = uv_lvl & 3;
    uv_sec += uv_sec == 3;
    uv_sec <<= bdm8;
    const int damping = a.damping + bdm8;
    const int bdmax = a.g.bdmax;  // 8-bit frames run with P = uint8_t, 10- and 12-bit with uint16_t
    // every CCSO neighbour (|dx| <= 2, |dy| <= 1) of the superblock's luma is
    // inside the frame: no clamping
    const bool cc_inner = x0 >= 2 && y0 >= 1 && x0 + 66 <= fw && y0 + 65 <= fh;
    __syncthreads();
    // ---- direction search: lane t & 63 owns 8x8 block (t & 7, (t >> 3) & 7)
    auto block_searched = [&](int b) -> bool {
        const int bx = (x0 >> 2) + (b & 7) * 2, by = (y0 >> 2) + (b >> 3) * 2;  // 4px units
        if (!active || !(y_pri || uv_pri) || bx >= a.g.bw || by >= a.g.bh) return false;
        return (noskip_b & (3u << (bx & 14))) != 0;
    };
    {
        const int b = t & 63, role = __builtin_amdgcn_readfirstlane(t >> 6);
        uint32_t c_lo = 0, c_hi = 0;
        if (block_searched(b)) {
            const int16_t *img = &ly[((b >> 3) * 8 + 2) * LY + (b & 7) * 8 + CX];
            if (role == 0)
                dir_costs<0>(img, LY, bdm8, c_lo, c_hi);
            else if (role == 1)
                dir_costs<1>(img, LY, bdm8, c_lo, c_hi);
            else if (role == 2)
                dir_costs<2>(img, LY, bdm8, c_lo, c_hi);
            else
                dir_costs<3>(img, LY, bdm8, c_lo, c_hi);
        }
        const int dlo = role == 0 ? 0 : role == 1 ? 2 : role == 2 ? 1 : 3;
        sm.dcost[dlo][b] = c_lo;
        sm.dcost[dlo + 4][b] = c_hi;
    }
    __syncthreads();
    if (t < 64) {
        int bx = (x0 >> 2) + (t & 7) * 2, by = (y0 >> 2) + (t >> 3) * 2;  // 4px units
        uint8_t st = 0, dir = 0;
        int16_t ypri = 0;
        if (active && bx < a.g.bw && by < a.g.bh) {
            if (noskip_b & (3u << (bx & 14))) {
                uint32_t var = 0;
                int d = 0;
                if (y_pri || uv_pri) {
                    // first maximum in direction order, as cdef_find_dir_bd
                    uint32_t best = sm.dcost[0][t];
#pragma unroll
                    for (int n = 1; n < 8; n++) {
                        const uint32_t c = sm.dcost[n][t];
                        if (c > best) {
                            best = c;
                            d = n;
                        }
                    }
                    var = (best - sm.dcost[d ^ 4][t]) >> 10;
                }
                // a lossless 4x4 in the block leaves that plane unfiltered
                // (cdef.rs:922-926, 984-1025)
                const int r = (t >> 3) * 2, sh = (t & 7) * 2;
                const bool y_ll = ((sm.ll[r] | sm.ll[r + 1]) >> sh) & 3;
                bool uv_ll;
                if constexpr (!kAny) {
                    uv_ll = (sm.ll[16 + (t >> 3)] >> (t & 7)) & 1;
                } else {
                    // the block's chroma 4x4 units: rows r >> ss_ver (two without vertical
                    // subsampling), columns sh >> ss_hor (two without horizontal)
                    const int cr = r >> ssy;
                    const uint32_t crow = sm.ll[16 + cr] | (ssy ? 0u : (uint32_t)sm.ll[16 + cr + 1]);
                    uv_ll = (crow >> (sh >> ssx)) & (ssx ? 1u : 3u);
                }
                if (y_pri) {
                    int adj = adjust_strength(y_pri, var);
                    if ((adj || y_sec) && !y_ll) st |= 1;
                    ypri = (int16_t)adj;
                } else if (y_sec && !y_ll) {
                    st |= 1;
                }
                if (uv_lvl && !uv_ll) st |= 2;
                dir = (uint8_t)d;
            }
        }
        b_state[t] = st;
        b_dir[t] = dir;
        b_ypri[t] = ypri;
    }
    __syncthreads();
    // ---- luma: lane t filters the column pair x = 2 (t & 31), 2 (t & 31) + 1
    // of the 8 rows of block row t >> 5 (one 8x8 block per lane)
    {
        const int x = 2 * (t & 31), by = 8 * (t >> 5);
        const int gx = x0 + x;
        const int bi = (by >> 3) * 8 + (x >> 3);
        const bool on = b_state[bi] & 1;
        const int pri = y_pri ? b_ypri[bi] : 0;
        const int dir = y_pri ? b_dir[bi] : 0;
        const int pri_tap = 4 - ((pri >> bdm8) & 1);
        const int pri_shift = pri ? dmax(0, damping - ulog2((uint32_t)pri)) : 0;
        const int sec_shift = y_sec ? damping - ulog2((uint32_t)y_sec) : 0;
        int op[2], os0[2], os1[2];
#pragma unroll
        for (int k = 0; k < 2; k++) {
            op[k] = sm.off[0][dir + 2][k];
            os0[k] = sm.off[0][dir + 4][k];
            os1[k] = sm.off[0][dir][k];
        }
        const bool cc0 = a.ccso_on && (ccso_pl & 1);
        if (gx < fw) {
#pragma unroll 2
            for (int yy = by; yy < by + 8; yy++) {
                const int gy = y0 + yy;
                if (gy >= fh) break;
                const int16_t *tp = &ly[(yy + 2) * LY + x + CX];
                int v0 = tp[0], v1 = tp[1];
                if (on) {
                    // the reference stores the result through a truncating cast to
                    // the sample type (Pixel::from_i32), before CCSO reads it back
                    const s16x2 r = cdef_px2(tp, op, os0, os1, pri, pri_tap, pri_shift, y_sec, sec_shift);
                    v0 = (int)(P)(uint16_t)r.x;
                    v1 = (int)(P)(uint16_t)r.y;
                }
                if (cc0 && !((sm.ll[yy >> 2] >> (x >> 2)) & 1)) {  // ccso_add skips lossless 4x4 units
                    const int8_t *lut = sm.ccso_lut[0];
                    int o0, o1;
                    if (cc_inner) {
                        ccso_pair_inner<1>(a.ccso[0], lut, tp, a.ccso[0].dy * LY + a.ccso[0].dx, o0, o1);
                    } else {
                        o0 = ccso_offset(a.ccso[0], lut, ly, gx, gy, x0, y0, fw, fh);
                        o1 = ccso_offset(a.ccso[0], lut, ly, gx + 1, gy, x0, y0, fw, fh);
                    }
                    v0 = iclip(v0 + o0, 0, bdmax);
                    v1 = iclip(v1 + o1, 0, bdmax);
                }
                store2(row_ptr<P>(a.dst[0], gy) + gx, v0, v1);
            }
        }
    }
    if constexpr (!kAny) {
        // ---- chroma: lane t filters the column pair 2 (t & 15), +1 of the 4 rows of
        // block row (t >> 4) & 7 of plane t >> 7 (one 4x4 block per lane)
        {
            const int pl = __builtin_amdgcn_readfirstlane(t >> 7), x = 2 * (t & 15), by = 4 * ((t >> 4) & 7);
            const int gx = (x0 >> 1) + x;
            const int bi = (by >> 2) * 8 + (x >> 2);
            const bool on = b_state[bi] & 2;
            const int dir = uv_pri ? b_dir[bi] : 0;
            const int pri_tap = 4 - ((uv_pri >> bdm8) & 1);
            const int pri_shift = uv_pri ? dmax(0, damping - 1 - ulog2((uint32_t)uv_pri)) : 0;
            const int sec_shift = uv_sec ? damping - 1 - ulog2((uint32_t)uv_sec) : 0;
            int op[2], os0[2], os1[2];
    #pragma unroll
            for (int k = 0; k < 2; k++) {
                op[k] = sm.off[1][dir + 2][k];
                os0[k] = sm.off[1][dir + 4][k];
                os1[k] = sm.off[1][dir][k];
            }
            const bool cc = a.ccso_on && ((ccso_pl >> (1 + pl)) & 1);
            if (a.has_chroma && gx < cw) {
    #pragma unroll 2
                for (int yy = by; yy < by + 4; yy++) {
                    const int gy = (y0 >> 1) + yy;
                    if (gy >= ch) break;
                    const int16_t *tp = &lc[pl][(yy + 2) * LC + x + CX];
                    int v0 = tp[0], v1 = tp[1];
                    if (on) {
                        const s16x2 r = cdef_px2(tp, op, os0, os1, uv_pri, pri_tap, pri_shift, uv_sec, sec_shift);
                        v0 = (int)(P)(uint16_t)r.x;
                        v1 = (int)(P)(uint16_t)r.y;
                    }
                    if (cc && !((sm.ll[16 + (yy >> 2)] >> (x >> 2)) & 1)) {
                        const CcsoDev &c = a.ccso[1 + pl];
                        const int8_t *lut = sm.ccso_lut[1 + pl];
                        int o0, o1;
                        if (cc_inner) {
                            ccso_pair_inner<2>(c, lut, &ly[(2 * yy + 2) * LY + 2 * x + CX], c.dy * LY + c.dx, o0, o1);
                        } else {
                            o0 = ccso_offset(c, lut, ly, gx << 1, gy << 1, x0, y0, fw, fh);
                            o1 = ccso_offset(c, lut, ly, (gx + 1) << 1, gy << 1, x0, y0, fw, fh);
                        }
                        v0 = iclip(v0 + o0, 0, bdmax);
                        v1 = iclip(v1 + o1, 0, bdmax);
                    }
                    store2(row_ptr<P>(a.dst[1 + pl], gy) + gx, v0, v1);
                }
            }
        }
    } else {
        // ---- chroma, in 32x32 sub-tiles of each plane's (64 >> ss_hor) x (64 >> ss_ver)
        // region (one for 4:2:0, two for 4:2:2, four for 4:4:4), each staged in lc in
        // turn: lane t filters the column pair 2 (t & 15), +1 of the 4 rows 4 ((t >> 4) & 7)..
        // of plane t >> 7, inside one luma 8x8 block
        const int nsx = 2 >> ssx, nsub = nsx * (2 >> ssy);
        // chroma direction from the luma direction: identity, or the 4:2:2 remap
        // (cdef.rs:605) {7, 0, 2, 4, 5, 6, 6, 6} as nibbles
        const uint32_t uvdir_map = (ssx && !ssy) ? 0x66654207u : 0x76543210u;
        for (int st = 0; st < nsub; st++) {
            const int sxo = (st % nsx) * 32, syo = (st / nsx) * 32;  // sub-tile origin in the region
            if (st > 0) {
                // once every lane is done with the previous sub-tile's LDS samples,
                // stage this one (4:2:2 / 4:4:4 only; loaded and stored group by group)
                __syncthreads();
                const int cx0 = (x0 >> ssx) + sxo, cy0 = (y0 >> ssy) + syo;
                for (int i = t; i < kNC; i += kThreads) {
                    const int cp = i >= LCR * kQC;
                    const int k = i - cp * LCR * kQC;
                    const int r = k / kQC, gx = cx0 - CX + (k - r * kQC) * 4, gy = cy0 - 2 + r;
                    uint2 v = make_uint2(kOutside, kOutside);
                    if (gx >= 0 && gx < cw && gy >= 0 && gy < ch) v = load4(row_ptr<const P>(a.src[1 + cp], gy) + gx);
                    ((uint2 *)&lc[0][0])[i] = v;
                }
                __syncthreads();
            }
            const int pl = __builtin_amdgcn_readfirstlane(t >> 7), x = 2 * (t & 15), by = 4 * ((t >> 4) & 7);
            const int cx = sxo + x, cyb = syo + by;  // chroma position in the superblock's region
            const int gx = (x0 >> ssx) + cx;
            const int bi = ((cyb << ssy) >> 3) * 8 + ((cx << ssx) >> 3);  // the luma 8x8 block
            const bool on = b_state[bi] & 2;
            const int dir = uv_pri ? (int)((uvdir_map >> (4 * b_dir[bi])) & 15) : 0;
            const int pri_tap = 4 - ((uv_pri >> bdm8) & 1);
            const int pri_shift = uv_pri ? dmax(0, damping - 1 - ulog2((uint32_t)uv_pri)) : 0;
            const int sec_shift = uv_sec ? damping - 1 - ulog2((uint32_t)uv_sec) : 0;
            int op[2], os0[2], os1[2];
    #pragma unroll
            for (int k = 0; k < 2; k++) {
                op[k] = sm.off[1][dir + 2][k];
                os0[k] = sm.off[1][dir + 4][k];
                os1[k] = sm.off[1][dir][k];
            }
            const bool cc = a.ccso_on && ((ccso_pl >> (1 + pl)) & 1);
            if (a.has_chroma && gx < cw) {
    #pragma unroll 2
                for (int yy = by; yy < by + 4; yy++) {
                    const int cy = syo + yy;
                    const int gy = (y0 >> ssy) + cy;
                    if (gy >= ch) break;
                    const int16_t *tp = &lc[pl][(yy + 2) * LC + x + CX];
                    int v0 = tp[0], v1 = tp[1];
                    if (on) {
                        const s16x2 r = cdef_px2(tp, op, os0, os1, uv_pri, pri_tap, pri_shift, uv_sec, sec_shift);
                        v0 = (int)(P)(uint16_t)r.x;
                        v1 = (int)(P)(uint16_t)r.y;
                    }
                    if (cc && !((sm.ll[16 + (cy >> 2)] >> (cx >> 2)) & 1)) {
                        const CcsoDev &c = a.ccso[1 + pl];
                        const int8_t *lut = sm.ccso_lut[1 + pl];
                        int o0, o1;
                        if (cc_inner) {
                            // collocated luma of the left sample; the right one's is 1 << ss_hor further
                            const int16_t *lp = &ly[((cy << ssy) + 2) * LY + (cx << ssx) + CX];
                            if (ssx)
                                ccso_pair_inner<2>(c, lut, lp, c.dy * LY + c.dx, o0, o1);
                            else
                                ccso_pair_inner<1>(c, lut, lp, c.dy * LY + c.dx, o0, o1);
                        } else {
                            o0 = ccso_offset(c, lut, ly, gx << ssx, gy << ssy, x0, y0, fw, fh);
                            o1 = ccso_offset(c, lut, ly, (gx + 1) << ssx, gy << ssy, x0, y0, fw, fh);
                        }
                        v0 = iclip(v0 + o0, 0, bdmax);
                        v1 = iclip(v1 + o1, 0, bdmax);
                    }
                    store2(row_ptr<P>(a.dst[1 + pl], gy) + gx, v0, v1);
                }
            }
        }
    }
}

// kAny instantiations for frames of any chroma subsampling (4:2:2 / 4:4:4),
// picked by the launchers
template <typename P, bool kAny>
__device__ __forceinline__ void cdef_body(const CdefArgs &a, int sbx, int sby, CdefSmem &sm) {
    CdefLoads L;
    cdef_load<P, kAny>(a, sbx, sby, L);
    cdef_store(L, sm);
    cdef_filter<P, kAny>(a, sbx, sby, L, sm);
}

template <typename P, bool kAny>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6, 8))) void cdef_kernel(CdefArgs a) {
    __shared__ CdefSmem sm;
    // XCD-contiguous superblock runs (neighbours share halo rows and columns in
    // L2), costliest first inside each run
    const int sbw = (a.g.bw * 4 + 63) / 64;
    const int slot = xcd_item(blockIdx.x, gridDim.x);
    const int sb = a.order ? a.order[slot] : slot;
    cdef_body<P, kAny>(a, sb % sbw, sb / sbw, sm);
}

// Batched launch: one item per 64x64 superblock of every frame of one bit depth.
template <typename P, bool kLoop, bool kAny>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6, 8))) void cdef_batch(const CdefArgs *__restrict__ args,
                                                                                        BatchSched s) {
    __shared__ CdefSmem sm;
    for_each_item<kLoop>(s, [&](int f, int t) {
        const CdefArgs &a = args[f];
        const int sbw = (a.g.bw * 4 + 63) >> 6;
        const int sby = t / sbw, sbx = t - sby * sbw;
        cdef_body<P, kAny>(a, sbx, sby, sm);
        __syncthreads();
    });
}

// Mixed launch: the superblocks of every bit-depth group's frames.
template <bool kLoop, bool kAny>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6, 8))) void cdef_mix(DevArgs3<CdefArgs> args,
                                                                                      MixSched s) {
    __shared__ CdefSmem sm;
    for_each_mixed_item<kLoop>(s, [&](int g, int f, int t) {
        const CdefArgs &a = args.p[g][f];
        const int sbw = (a.g.bw * 4 + 63) >> 6;
        const int sby = t / sbw, sbx = t - sby * sbw;
        if (g == 0)
            cdef_body<uint8_t, kAny>(a, sbx, sby, sm);
        else
            cdef_body<uint16_t, kAny>(a, sbx, sby, sm);
        __syncthreads();
    });
}

void launch_cdef_mix(DevArgs3<CdefArgs> args, const MixSched &s, int grid, bool any, hipStream_t st) {
    if (s.first[3] <= 0) return;
    const bool loop = grid < s.first[3];
    if (any) {
        if (loop)
            R2G_LAUNCH((cdef_mix<true, true>), dim3(grid), dim3(kThreads), 0, st, args, s);
        else
            R2G_LAUNCH((cdef_mix<false, true>), dim3(grid), dim3(kThreads), 0, st, args, s);
    } else {
        if (loop)
            R2G_LAUNCH((cdef_mix<true, false>), dim3(grid), dim3(kThreads), 0, st, args, s);
        else
            R2G_LAUNCH((cdef_mix<false, false>), dim3(grid), dim3(kThreads), 0, st, args, s);
    }
}

int cdef_items(const CdefArgs &a) { return ((a.g.bw * 4 + 63) >> 6) * ((a.g.bh * 4 + 63) >> 6); }

// Per device: the direction offsets as (dy, dx) and the CCSO offset table.
void cdef_ccso_lut(CcsoDev &d, const uint8_t filter_off[64], int scale_idx) {
    // ccso_add (ccso.rs:215-254): index -> packed nibble of filter_off -> CCSO_OFFSET
    for (int idx = 0; idx < 256; idx++) {
        const int packed = (idx >> 1) < 64 ? filter_off[idx >> 1] : 0;
        d.lut[idx] = R2T_CCSO_OFFSET[scale_idx & 3][7 & (packed >> (4 * (idx & 1)))];
    }
}

void cdef_tap_offsets(CdefArgs &a) {
    for (int d = 0; d < 12; d++)
        for (int k = 0; k < 2; k++) {
            const int o = R2T_CDEF_DIRECTIONS[d][k];  // dy * 12 + dx with |dx| <= 2
            const int dx = ((o + 2) % 12 + 12) % 12 - 2, dy = (o - dx) / 12;
            a.off[0][d][k] = (int16_t)(dy * LY + dx);
            a.off[1][d][k] = (int16_t)(dy * LC + dx);
        }
}

void launch_cdef_batch(const CdefArgs *d_args, const BatchSched &s, int grid, int bitdepth, bool any,
                       hipStream_t st) {
    if (s.total <= 0) return;
    const bool loop = grid < s.total;
#define R2G_CB(P, L)                                                                                       \
    do {                                                                                                   \
        if (any)                                                                                           \
            R2G_LAUNCH((cdef_batch<P, L, true>), dim3(grid), dim3(kThreads), 0, st, d_args, s);          \
        else                                                                                               \
            R2G_LAUNCH((cdef_batch<P, L, false>), dim3(grid), dim3(kThreads), 0, st, d_args, s);         \
    } while (0)
    if (bitdepth > 8) {
        if (loop)
            R2G_CB(uint16_t, true);
        else
            R2G_CB(uint16_t, false);
    } else {
        if (loop)
            R2G_CB(uint8_t, true);
        else
            R2G_CB(uint8_t, false);
    }
#undef R2G_CB
}

void launch_cdef(const CdefArgs &a, int bitdepth, hipStream_t s) {
    dim3 grid(((a.g.bw * 4 + 63) / 64) * ((a.g.bh * 4 + 63) / 64));
    const bool any = a.uvm != nullptr;
    if (bitdepth > 8) {
        if (any)
            R2G_LAUNCH((cdef_kernel<uint16_t, true>), grid, dim3(kThreads), 0, s, a);
        else
            R2G_LAUNCH((cdef_kernel<uint16_t, false>), grid, dim3(kThreads), 0, s, a);
    } else {
        if (any)
            R2G_LAUNCH((cdef_kernel<uint8_t, true>), grid, dim3(kThreads), 0, s, a);
        else
            R2G_LAUNCH((cdef_kernel<uint8_t, false>), grid, dim3(kThreads), 0, s, a);
    }
}

}  // namespace r2g
