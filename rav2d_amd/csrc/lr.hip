// Loop restoration on gfx950: NS-Wiener (single luma, single chroma with
// cross-component luma taps, multi-class), PC-Wiener and GDF.  One workgroup
// per (plane, stripe, 64-column unit) work item; F3 (+ F2 boundary rows) -> F4.
//
// Arithmetic follows crates/rav2d/src/looprestoration.rs: ns_wiener_single_y
// (:524-713), wiener_multi with get_class_lut_idx / get_qval_given_tskip
// (:715-1059, 195-271, 40-49), ns_wiener_single_uv with backup_row_luma
// (:1145-1430, 408-522), gdf_prep with compute_gradient_row (:1432-1643,
// 114-192) and gdf_add (:273-329).
//
// The reference slides a window of row buffers down each stripe: rows inside
// the stripe come from the CDEF output, the two rows beyond each stripe
// boundary from the deblocked-line store, and further rows repeat the last
// available one; columns beyond a unit come from its neighbours (pre-LR) or
// repeat the edge column at the frame's left/right border.  Here that virtual
// image is staged once per work item in LDS and every output sample reads it
// directly.
#include <type_traits>

#include "internal.h"
#include "tables.h"

namespace r2g {

__constant__ uint16_t kPcNorm[4] = {3739, 3273, 3074, 7};
__constant__ int16_t kModeW[4][3] = {{-527, 15325, 321}, {26436, -17705, 17905}, {366, -147, -194}, {202, -267, -179}};
__constant__ int16_t kModeOff[4] = {-547, -21565, -573, -680};

constexpr int kThreads = 256;
constexpr int E = 6;               // staged halo (GDF reach)
constexpr int VC = 8;              // column of item sample 0 in a staged row
constexpr int VW = 64 + 2 * VC;    // staged row: item columns [-8, 72), loaded 4 samples at a time
constexpr int VQ = VW / 4;
constexpr int VH = 64 + 2 * E;     // staged height

__device__ __forceinline__ int qval_tskip(int qstep, int tskip, int i, int bdm8) {
    qstep = (qstep + ((1 << bdm8) >> 1)) >> bdm8;
    int prod = (tskip * qstep + 128) >> 8;
    int q = kModeW[i][0] * (tskip << 5) + kModeW[i][1] * qstep + kModeW[i][2] * prod;
    q = apply_sign((abs(q) + (1 << 12)) >> 13, q);
    return 255 * (kModeOff[i] + q);
}

// Compile-time tap tables (offsets fold into LDS immediate offsets).
struct Tap {
    int8_t dy, dx;
};
constexpr Tap kNsYT[16] = {{1, 0}, {0, 1}, {2, 0}, {0, 2}, {1, 1}, {-1, 1}, {2, 1}, {2, -1},
                           {1, 2}, {1, -2}, {3, 0}, {0, 3}, {4, 0}, {0, 4}, {3, 3}, {3, -3}};
constexpr Tap kPcT[12] = {{1, 0}, {0, 1}, {2, 0}, {0, 2}, {1, 1}, {-1, 1}, {2, 1}, {2, -1}, {1, 2}, {1, -2}, {3, 0}, {0, 3}};
constexpr Tap kNsUvT[6] = {{1, 0}, {0, 1}, {1, 1}, {-1, 1}, {2, 0}, {0, 2}};
constexpr Tap kNsUvYT[12] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, -1}, {-1, 1}, {1, -1}, {2, 0}, {-2, 0}, {0, 2}, {0, -2}};
constexpr Tap kGdfT[18] = {{6, 0}, {5, 0}, {4, 0}, {3, 0}, {2, 1}, {2, 0}, {2, -1}, {1, 2}, {1, 1},
                           {1, 0}, {1, -1}, {1, -2}, {0, 6}, {0, 5}, {0, 4}, {0, 3}, {0, 2}, {0, 1}};

constexpr int LS = 64 + 4;  // stride of the downsampled luma rows (2-sample halo)

typedef short s16x2 __attribute__((ext_vector_type(2)));

// Narrows a value known to fit 16 bits (sums and differences of at most four
// 8/10-bit samples, clamped GDF features) so that products with filter taps
// compile to 24-bit multiplies.
__device__ __forceinline__ int s16(int v) { return (int)(int16_t)v; }
__device__ __forceinline__ uint32_t pack16(int lo, int hi) { return (uint32_t)(uint16_t)lo | ((uint32_t)(uint16_t)hi << 16); }

// Wiener taps of a 2x2 quad two at a time: the pair of samples of taps 2i and
// 2i+1 (sums of two samples for symmetric taps, at most 2046) in packed
// 16-bit lanes, dotted with the packed 16-bit filter pair f[i] into each
// pixel's 32-bit sum.  Stage by stage across the four pixels, so that no
// packed operation waits on the one before it.  pp[p] points at pixel p in a
// sample array of row stride S.
template <const Tap *T, int NP, int S>
__device__ __forceinline__ void sym_pairs(const uint16_t *const (&pp)[4], const uint32_t *f, int (&s)[4]) {
#pragma unroll
    for (int i = 0; i < NP; i++) {
        s16x2 x[4];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const s16x2 A = {(short)pp[p][-T[2 * i].dy * S - T[2 * i].dx], (short)pp[p][-T[2 * i + 1].dy * S - T[2 * i + 1].dx]};
            const s16x2 B = {(short)pp[p][T[2 * i].dy * S + T[2 * i].dx], (short)pp[p][T[2 * i + 1].dy * S + T[2 * i + 1].dx]};
            x[p] = A + B;
        }
#pragma unroll
        for (int p = 0; p < 4; p++) s[p] = __builtin_amdgcn_sdot2(x[p], __builtin_bit_cast(s16x2, f[i]), s[p], false);
    }
}
// the same for one-sided taps (the cross-component luma taps of chroma)
template <const Tap *T, int NP, int S>
__device__ __forceinline__ void tap_pairs(const uint16_t *const (&pp)[4], const uint32_t *f, int (&s)[4]) {
#pragma unroll
    for (int i = 0; i < NP; i++) {
        s16x2 x[4];
#pragma unroll
        for (int p = 0; p < 4; p++)
            x[p] = s16x2{(short)pp[p][T[2 * i].dy * S + T[2 * i].dx], (short)pp[p][T[2 * i + 1].dy * S + T[2 * i + 1].dx]};
#pragma unroll
        for (int p = 0; p < 4; p++) s[p] = __builtin_amdgcn_sdot2(x[p], __builtin_bit_cast(s16x2, f[i]), s[p], false);
    }
}
constexpr int CS = 33;      // stride of the GDF gradient cells

// GDF error-table index of each pixel's three weighted sums (gdf_prep,
// looprestoration.rs:1620-1630): sum * G (G = 8 intra, 5 inter) rounded half
// away from zero at 2^15, clipped to [-G, G-1], offset by G and combined in
// base 2G.  The rounding is taken as a floor, (s + 2^14 - [s < 0]) >> 15,
// with the factor 8 folded into the shift; the constant part of the offsets,
// G (4G^2 + 2G + 1), moves into the table base (kGdfErrBase*).
template <int G> __device__ __forceinline__ void gdf_index(const int (&iv)[4][3], int (&idx)[4]) {
#pragma unroll
    for (int p = 0; p < 4; p++) {
        int v[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int x = iv[p][k];
            int q;
            if constexpr (G == 8) {
                q = (x + (1 << 11) + (x >> 31)) >> 12;
            } else {
                const int s = (x << 2) + x;
                q = (s + (1 << 14) + (s >> 31)) >> 15;
            }
            v[k] = dmin(dmax(q, -G), G - 1);
        }
        idx[p] = __mul24(v[0], 4 * G * G) + __mul24(v[1], 2 * G) + v[2];
    }
}
constexpr int kGdfErrBase8 = 8 * (4 * 64 + 2 * 8 + 1), kGdfErrBase5 = 5 * (4 * 25 + 2 * 5 + 1);

struct LrSmem {
    alignas(16) uint16_t V[VH * VW];  // virtual image, rows [-E, 64 + E), columns [-VC, 64 + VC)
    union {
        uint16_t Ld[(64 + 4) * LS];  // the luma taps of a chroma stripe (chroma items): rows -2 .. h + 1
        struct {
            uint16_t cells[33 * CS * 4];     // GDF gradient cells (luma items)
            alignas(16) int8_t s_err[4096];  // GDF error table slice of (ref, qp)
        };
    };
    uint8_t cls[16 * 16];            // class of each 4x4 block
    alignas(16) uint16_t s_alpha[88];
    alignas(16) int16_t s_weight[264];
    alignas(16) uint32_t s_pcp[64 * 8];  // PC-Wiener filters of the frame's wiener_idx, packed tap pairs
    alignas(16) uint32_t s_mfp[16 * 8];  // multi-class frame filters of the plane, packed tap pairs
    int32_t s_mF[16];                    // their centre weights 128 - 2 sum(f)
    uint32_t apk[4 * 9];                // GDF alpha of taps (2i, 2i+1) per class, packed 16-bit pairs
    uint32_t wpk[4 * 27];               // GDF weights likewise: [class][weight set][pair]
    int16_t s_bias[4];                  // GDF bias of the three weight sets for (ref, qp)
    LrItem it;
};

template <typename P, int BD, typename Args>
__device__ __forceinline__ void lr_body(Args &a, int item, LrSmem &sm) {
    uint16_t *V = sm.V, *Ld = sm.Ld, *cells = sm.cells;
    uint8_t *cls = sm.cls;
    LrItem &it = sm.it;
    const int t = threadIdx.x;
    if (t < sizeof(LrItem) / 4) ((uint32_t *)&it)[t] = ((const uint32_t *)&a.items[item])[t];
    __syncthreads();
    // item fields are uniform over the workgroup
    const int pl = __builtin_amdgcn_readfirstlane(it.plane), x0 = __builtin_amdgcn_readfirstlane(it.x0),
              w = __builtin_amdgcn_readfirstlane(it.w), y0 = __builtin_amdgcn_readfirstlane(it.y0),
              h = __builtin_amdgcn_readfirstlane(it.h);
    const int edges = __builtin_amdgcn_readfirstlane(it.edges);
    const int kind = __builtin_amdgcn_readfirstlane(it.kind);
    const bool gdf = __builtin_amdgcn_readfirstlane(it.gdf_on) != 0 && pl == 0;  // GDF is luma-only
    auto &f3 = a.f3[pl];
    auto &f2 = a.f2[pl];
    auto &dst = a.dst[pl];
    // one instantiation per bit depth (8-bit frames with P = uint8_t, 10- and
    // 12-bit with uint16_t), so clip bounds and shifts are constants
    constexpr int bd = BD, bdm8 = bd - 8, bdmax = (1 << bd) - 1;
    static_assert(sizeof(P) == (BD > 8 ? 2 : 1), "sample type of the bit depth");
    // every buffer of a plane has the same row stride; rows and strides are
    // below 2^24, so row offsets are 24-bit products
    const uint32_t stride = (uint32_t)f3.stride;
    uint8_t *const dbase = dst.ptr + (uint64_t)(uint32_t)y0 * stride + x0 * sizeof(P);
    using Q = typename std::conditional<sizeof(P) == 1, uint32_t, uint2>::type;  // 4 samples
    if (kind == LR_KIND_NONE && !gdf) {
        // plane copy, 4 samples per access (x0 and w are multiples of 4): units of
        // planes LR leaves alone, or blocks of up to 256 x 64 of planes it does
        // not touch; per round, loads from clamped indices first, then the stores
        constexpr int CP = 8;
        const int n4 = w >> 2, n = h * n4;
        const uint8_t *src = f3.ptr + (uint64_t)(uint32_t)y0 * stride + x0 * sizeof(P);
        uint8_t *dp = dst.ptr + (uint64_t)(uint32_t)y0 * stride + x0 * sizeof(P);
        // element t + j * kThreads of the h x n4 grid, stepped without divisions
        const int dq = kThreads / n4, dm = kThreads - dq * n4;
        int r = t / n4, c = t - (t / n4) * n4;
        for (int b = 0; b < n; b += CP * kThreads) {
            Q v[CP];
            int rs[CP], cs[CP];
#pragma unroll
            for (int j = 0; j < CP; j++) {
                rs[j] = r;
                cs[j] = c;
                v[j] = *(const Q *)(src + __umul24(dmin(r, h - 1), stride) + c * sizeof(Q));
                r += dq;
                c += dm;
                if (c >= n4) {
                    c -= n4;
                    r++;
                }
            }
#pragma unroll
            for (int j = 0; j < CP; j++)
                if (rs[j] < h) *(Q *)(dp + __umul24(rs[j], stride) + cs[j] * sizeof(Q)) = v[j];
        }
        return;
    }
    // ---- staging.  Each lane issues all of its global loads before its first
    // LDS store, so the load latencies overlap instead of adding up.
    // Virtual image V: rows [-E, h+E), item columns [-8, 72) in groups of 4
    // samples read at once (x0 and plane widths are multiples of 4).  Groups
    // are read from clamped, always valid addresses; the columns the
    // reference replicates at the frame's left/right edge are patched after.
    // Rows: inside the stripe from F3; across a stripe boundary the two
    // deblocked rows beyond it (F2), further rows repeating the second; at the
    // frame's top/bottom the stripe's first/last row repeats.
    // lane t < RPP * VQ owns column group t % VQ of staged rows t / VQ + RPP j:
    // the column clamp is computed once, rows step by a constant
    constexpr int RPP = kThreads / VQ, PERQ = (VH + RPP - 1) / RPP;
    const int pw = f3.w;
    const bool top = edges & LRE_TOP, bottom = edges & LRE_BOTTOM;
    const int cq = t % VQ, r0 = t / VQ;
    const bool lane_on = t < RPP * VQ;
    const int col = iclip(x0 + cq * 4 - VC, 0, pw - 4) * (int)sizeof(P);
    Q qv[PERQ];
#pragma unroll
    for (int j = 0; j < PERQ; j++) {
        const int r = dmin(r0 + RPP * j, h + 2 * E - 1) - E;
        const bool from2 = (r < 0 && top) || (r >= h && bottom);
        const int yy = y0 + iclip(r, from2 ? -2 : 0, from2 ? h + 1 : h - 1);
        const uint8_t *row = (from2 ? f2.ptr : f3.ptr) + __umul24(yy, stride);
        qv[j] = *(const Q *)(row + col);
    }
    // tables of this item's filter kinds, 16 (or 8) bytes per lane
    uint4 pcv = {0, 0, 0, 0}, alv = {0, 0, 0, 0}, wv = {0, 0, 0, 0}, erv = {0, 0, 0, 0};
    uint2 erv2 = {0, 0};
    int biasv = 0;
    if (kind == LR_KIND_PC && t < 64 * 8 * 4 / 16) pcv = ((const uint4 *)(a.t.pc_pairs + a.widx * 64 * 8))[t];
    if (gdf) {
        if (t < 88 * 2 / 16) alv = ((const uint4 *)(a.t.gdf_alpha + a.gdf_ref * 528 + a.gdf_qp * 88))[t];
        if (t < 264 * 2 / 16) wv = ((const uint4 *)(a.t.gdf_weight + a.gdf_ref * 1584 + a.gdf_qp * 264))[t];
        if (t < 3) biasv = a.t.gdf_bias[(a.gdf_ref * 6 + a.gdf_qp) * 3 + t];
        if (a.gdf_ref == 0)
            erv = ((const uint4 *)(a.t.gdf_intra + a.gdf_qp * 4096))[t];  // 4096 B: one uint4 per lane
        else if (t < 1000 / 8)
            erv2 = ((const uint2 *)(a.t.gdf_inter + (a.gdf_ref - 1) * 6000 + a.gdf_qp * 1000))[t];
    }
    const bool uv_single = kind == LR_KIND_SINGLE && pl != 0;
    // 4:2:0 chroma averages luma row pairs, staged with the image below;
    // 4:2:2 / 4:4:4 chroma (no vertical subsampling) is staged after it
    const bool uv_ss_ver = a.g.ss_ver != 0;
    // lane t < LRP * LS owns column t % LS of the rows t / LS + LRP j
    constexpr int LRP = kThreads / LS, PERL = (32 + 4 + LRP - 1) / LRP;
    const int lc = t % LS, lr0 = t / LS;
    const bool llane = t < LRP * LS;
    uint16_t lvals[PERL];
    if (uv_single && uv_ss_ver) {
        // downsampled luma rows of the chroma stripe (cfl_ds_filter_index):
        // luma rows 2y, 2y + 1 inside the stripe (F3); across a stripe boundary
        // the deblocked rows (F2): above -2 -> (2y0-2, 2y0-2), -1 -> (2y0-2, 2y0-1),
        // below h -> (2(y0+h), +1), h+1 -> (2(y0+h)+1, 2(y0+h)+1)
        auto &l3 = a.f3[0];
        auto &l2 = a.f2[0];
        const uint32_t lstride = (uint32_t)l3.stride;
        // cfl_ds 0: 2x2 average, 1: vertical pair average, 2: top-left sample
        const int ds = a.cfl_ds, mb = ds <= 1 ? -1 : 0, m1 = ds == 0 ? -1 : 0, sh = ds == 0 ? 2 : (ds == 1 ? 1 : 0);
        using Q2 = typename std::conditional<sizeof(P) == 1, uint16_t, uint32_t>::type;  // 2 samples
        const int c = lc - 2;
        const int xs = (c < 0 && !(edges & LRE_LEFT)) ? 0 : ((c >= w && !(edges & LRE_RIGHT)) ? w - 1 : dmin(c, w + 1));
        const int lx = 2 * (x0 + xs) * (int)sizeof(P);
#pragma unroll
        for (int j = 0; j < PERL; j++) {
            const int gr = lr0 + LRP * j;
            const int r = dmin(gr, h + 3) - 2;
            const int rr = iclip(r, top ? -2 : 0, bottom ? h + 1 : h - 1);
            const bool from2 = rr < 0 || rr >= h, single = rr == -2 || rr == h + 1;
            const int ya = 2 * (y0 + rr) + (rr == -2 ? 2 : (rr == h + 1 ? -1 : 0));
            const uint8_t *ra = (from2 ? l2.ptr : l3.ptr) + __umul24(ya, lstride);
            const uint8_t *rb = single ? ra : ra + lstride;
            const Q2 qa = *(const Q2 *)(ra + lx), qb = *(const Q2 *)(rb + lx);
            constexpr int SB = 8 * sizeof(P);
            const int a0 = (int)(qa & ((1u << SB) - 1)), a1 = (int)(qa >> SB);
            const int b0 = (int)(qb & ((1u << SB) - 1)), b1 = (int)(qb >> SB);
            const int v = (a0 + (b0 & mb) + ((a1 + b1) & m1)) >> sh;
            lvals[j] = (gr < h + 4 && c < w + 2) ? (uint16_t)v : 0;
        }
    }
#pragma unroll
    for (int j = 0; j < PERQ; j++) {
        const int rr = r0 + RPP * j;
        if (lane_on && rr < h + 2 * E) {
            const int i = rr * VQ + cq;
            uint2 u;
            if constexpr (sizeof(P) == 1) {
                const uint32_t b = (uint32_t)qv[j];
                u.x = (b & 0xff) | ((b << 8) & 0xff0000);
                u.y = ((b >> 16) & 0xff) | ((b >> 8) & 0xff0000);
            } else {
                u = *(const uint2 *)&qv[j];
            }
            *(uint2 *)&V[i * 4] = u;
        }
    }
    if (uv_single && uv_ss_ver) {
#pragma unroll
        for (int j = 0; j < PERL; j++) {
            const int gr = lr0 + LRP * j;
            if (llane && gr < h + 4) Ld[gr * LS + lc] = lvals[j];
        }
    }
    if (uv_single && !uv_ss_ver) {
        // backup_row_luma with ss_ver = 0 copies luma rows (ns_wiener_single_uv,
        // looprestoration.rs:408-425, 1145-1430): chroma (r, c) reads luma row
        // y0 + r at column (x0 + c) << ss_hor; the unit's luma row extends 4
        // samples each side, replicating its first / last luma column at a
        // frame edge.  Above the stripe both rows read the line store's row 0
        // (the deblocked row y0 - 2), below it rows 6 and 7; no stripe beyond
        // the frame: the first / last stripe row repeats.  The store is read at
        // column 2 x0 + (c << ss_hor) whatever the subsampling (llpf_off = x * 2,
        // :1957-1959): for 4:4:4 its columns past the frame width are zeros up to
        // the store's row length and then continue in its next row.
        auto &l3 = a.f3[0];
        auto &l2 = a.f2[0];
        const int ssx = a.g.ss_hor, lw = w << ssx, lx0 = x0 << ssx;
        const int W = l3.w, S = a.lstore_stride;
        for (int i = t; i < (h + 4) * LS; i += kThreads) {
            const int gr = i / LS, c = i - gr * LS - 2;
            uint16_t v = 0;
            if (c < w + 2) {
                const int r = gr - 2;
                // luma offset of the tap in the unit's extended row
                const int o = c << ssx;
                const int oc = (o < 0 && !(edges & LRE_LEFT)) ? 0 : ((o >= lw && !(edges & LRE_RIGHT)) ? lw - 1 : o);
                int k = -1;  // line-store row slot: 0 above, 2 / 3 below
                int row = y0 + r;
                if (r < 0) {
                    if (top) k = 0; else row = y0;
                } else if (r >= h) {
                    if (bottom) k = r == h ? 2 : 3; else row = y0 + h - 1;
                }
                if (k < 0) {
                    v = (uint16_t)row_ptr<const P>(l3, row)[lx0 + oc];
                } else {
                    int sc = 2 * x0 + oc, src = it.lsrc[k];
                    if (sc >= S) {
                        sc -= S;
                        src = it.lsrc[k + 1];  // the store's next row
                    } else if (sc >= W) {
                        src = -1;  // past the frame width: never written
                    }
                    if (src >= 0) v = (uint16_t)row_ptr<const P>(l2, src)[sc];
                }
            }
            Ld[i] = v;
        }
    }
    if (kind == LR_KIND_PC && t < 64 * 8 * 4 / 16) ((uint4 *)sm.s_pcp)[t] = pcv;
    if (kind == LR_KIND_MULTI && t < 16 * 8) {
        // tap pairs of class t / 8; lane 0 of each class also sums its taps
        const auto &f = a.mfilt[pl][t >> 3];
        const int i = t & 7;
        sm.s_mfp[t] = pack16(f[2 * i], f[2 * i + 1]);
        if (i == 0) {
            int sum = 0;
#pragma unroll
            for (int k = 0; k < 16; k++) sum += f[k];
            sm.s_mF[t >> 3] = 128 - 2 * sum;
        }
    }
    if (gdf) {
        if (t < 88 * 2 / 16) ((uint4 *)sm.s_alpha)[t] = alv;
        if (t < 264 * 2 / 16) ((uint4 *)sm.s_weight)[t] = wv;
        if (t < 3) sm.s_bias[t] = (int16_t)biasv;
        if (a.gdf_ref == 0)
            ((uint4 *)sm.s_err)[t] = erv;
        else if (t < 1000 / 8)
            ((uint2 *)sm.s_err)[t] = erv2;
    }
    if (!(edges & LRE_LEFT) || !(edges & LRE_RIGHT)) {
        // frame edge: columns -6..-1 repeat column 0, columns w..w+5 repeat w-1
        __syncthreads();
        for (int i = t; i < (h + 2 * E) * 12; i += kThreads) {
            const int r = i / 12, k = i - r * 12;
            uint16_t *vr = &V[r * VW + VC];
            if (k < 6) {
                if (!(edges & LRE_LEFT)) vr[k - 6] = vr[0];
            } else if (!(edges & LRE_RIGHT)) {
                vr[w + k - 6] = vr[w - 1];
            }
        }
    }
    __syncthreads();
    // samples of V relative to the centre pointer
#define VAT(p, dy, dx) ((int)(p)[(dy) * VW + (dx)])
    // GDF works in the 10-bit domain: 8-bit differences are shifted up by 2,
    // 12-bit samples down by 2 before any gradient or difference (gdf_prep,
    // looprestoration.rs:1450-1453)
    constexpr int up = BD == 8 ? 2 : 0, down = BD == 12 ? 2 : 0;
    // ---- GDF gradient cells: cell (j,k) covers centre columns 2j-1,2j and rows 2k-1,2k
    const int ncx = w / 2 + 1, ncy = h / 2 + 1;
    if (gdf) {
        // tap-pair tables for the packed feature loop
        if (t < 4 * 9) {
            const int c = t / 9, q2 = t - c * 9;
            sm.apk[t] = (uint32_t)sm.s_alpha[(2 * q2) * 4 + c] | ((uint32_t)sm.s_alpha[(2 * q2 + 1) * 4 + c] << 16);
        } else if (t < 4 * 9 + 4 * 27) {
            const int i = t - 4 * 9, c = i / 27, r = i - c * 27, k = r / 9, q2 = r - k * 9;
            sm.wpk[i] = (uint32_t)(uint16_t)sm.s_weight[k * 88 + (2 * q2) * 4 + c] |
                        ((uint32_t)(uint16_t)sm.s_weight[k * 88 + (2 * q2 + 1) * 4 + c] << 16);
        }
        // lane t: cell column t % 33, rows t / 33, + 7, ... (ncx, ncy <= 33)
        const int j = t % 33;
        if (t < 7 * 33 && j < ncx) {
            for (int k = t / 33; k < ncy; k += 7) {
                // rows 2k-2 .. 2k+1, columns 2j-2 .. 2j+1 as aligned 16-bit pairs
                uint32_t wv[4][2];
                const uint32_t *w0 = (const uint32_t *)&V[(2 * k - 2 + E) * VW + 2 * j - 2 + VC];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    wv[r][0] = w0[r * (VW / 2)];
                    wv[r][1] = w0[r * (VW / 2) + 1];
                }
                auto S = [&](int r, int c) -> uint32_t {
                    return ((wv[r][c >> 1] >> (16 * (c & 1))) & 0xffff) >> down;
                };
                auto sad = [](uint32_t x, uint32_t y, uint32_t acc) { return (x > y ? x - y : y - x) + acc; };
                uint32_t g0 = 0, g1 = 0, g2 = 0, g3 = 0;
#pragma unroll
                for (int r = 1; r <= 2; r++)
#pragma unroll
                    for (int c = 1; c <= 2; c++) {
                        const uint32_t b2 = 2 * S(r, c);
                        g0 = sad(S(r - 1, c) + S(r + 1, c), b2, g0);
                        g1 = sad(S(r, c - 1) + S(r, c + 1), b2, g1);
                        g2 = sad(S(r - 1, c - 1) + S(r + 1, c + 1), b2, g2);
                        g3 = sad(S(r + 1, c - 1) + S(r - 1, c + 1), b2, g3);
                    }
                uint32_t *cp = (uint32_t *)&cells[(k * CS + j) * 4];
                cp[0] = g0 | (g1 << 16);
                cp[1] = g2 | (g3 << 16);
            }
        }
    }
    // ---- classes of the multi-class filters (one per 4x4 block)
    const int bw4 = w >> 2, bh4 = h >> 2;
    if (kind == LR_KIND_MULTI || kind == LR_KIND_PC) {
        const uint8_t *cmap = kind == LR_KIND_PC ? a.t.cls_pc + a.widx * 4096 : a.t.cls_ns + (a.widx * 7 + a.ns_idx) * 4096;
        // a unit has at most 16 x 16 blocks: one per lane
        if (t < bw4 * bh4) {
            const int by = t / bw4, bx = t - by * bw4;
            // the block's window, rows by*4-2 .. by*4+5 and columns bx*4-2 ..
            // bx*4+5, as four aligned 16-bit pairs per row, all read before use
            uint32_t wv[8][4];
            const uint32_t *w0 = (const uint32_t *)&V[(by * 4 - 2 + E) * VW + bx * 4 - 2 + VC];
#pragma unroll
            for (int r = 0; r < 8; r++)
#pragma unroll
                for (int c = 0; c < 4; c++) wv[r][c] = w0[r * (VW / 2) + c];
            auto S = [&](int r, int c) -> uint32_t { return (wv[r][c >> 1] >> (16 * (c & 1))) & 0xffff; };
            // |up + down - 2 m| summed with v_sad_u32 (sums of two samples <= 2046)
            auto sad = [](uint32_t x, uint32_t y, uint32_t acc) { return (x > y ? x - y : y - x) + acc; };
            uint32_t f0 = 0, f1 = 0, f2v = 0;
#pragma unroll
            for (int r = 1; r <= 6; r++)
#pragma unroll
                for (int c = 1; c <= 6; c++) {
                    const uint32_t m2 = 2 * S(r, c);
                    f0 = sad(S(r - 1, c) + S(r + 1, c), m2, f0);
                    f1 = sad(S(r - 1, c + 1) + S(r + 1, c - 1), m2, f1);
                    f2v = sad(S(r - 1, c - 1) + S(r + 1, c + 1), m2, f2v);
                }
            int s = 0;
            const int npx[3] = {16, 4, 1};
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int e = (dy != 0) + (dx != 0);
                    int fx = iclip((bx & 15) + dx, 0, 15), fy = iclip(by + dy, 0, bh4 - 1);
                    s += npx[e] * (((it.noskip[fy] >> fx) & 1) == 0);
                }
            int rnd = (1 << bdm8) >> 1;
            int f[3] = {(int)f0, (int)f1, (int)f2v};
#pragma unroll
            for (int k = 0; k < 3; k++) f[k] = (int)((uint32_t)f[k] * kPcNorm[k] + (uint32_t)rnd) >> bdm8;
            s *= kPcNorm[3];
            int q = (dmax(0, qval_tskip(a.base_q, s, 0, bdm8)) + (1 << 13)) >> 14;
            q = dmin(q, 255) >> 5;
            int li = q << 9;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                q = (dmax(0, f[k] + qval_tskip(a.base_q, s, k + 1, bdm8)) + (1 << 13)) >> 14;
                q = dmin(q, 255) >> 5;
                li |= q << (3 * (2 - k));
            }
            cls[by * 16 + bx] = cmap[li];
        }
    }
    __syncthreads();
    // ---- outputs, one 2x2 quad per lane: a quad shares its 4x4 block (class,
    // lossless flag) and its GDF cell sums (class c and gradient features)
    const bool intra = a.gdf_ref == 0;
    int gbias[3] = {0, 0, 0};
    if (gdf) {
#pragma unroll
        for (int k = 0; k < 3; k++) gbias[k] = __builtin_amdgcn_readfirstlane(sm.s_bias[k]);
    }
    // The NS-Wiener sum m << 7 + sum_k f_k (A_k + B_k - 2m) is evaluated as
    // m (128 - 2 sum_k f_k) + sum_k f_k (A_k + B_k): the same integer.  The
    // single filter's packed tap pairs and centre weights are uniform.
    uint32_t fs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int fc = 128, fl = 0;  // centre weight; minus the weight of the co-located luma sample (chroma)
    if (kind == LR_KIND_SINGLE) {
        const int nc = pl == 0 ? 16 : 6;
        int sc = 0, sl = 0;
#pragma unroll
        for (int k = 0; k < 18; k++) {
            const int f = it.filt[k];
            if (k < nc)
                sc += f;
            else
                sl += f;
        }
        fc = __builtin_amdgcn_readfirstlane(128 - 2 * sc);
        fl = __builtin_amdgcn_readfirstlane(sl);
#pragma unroll
        for (int i = 0; i < 9; i++) fs[i] = __builtin_amdgcn_readfirstlane(pack16(it.filt[2 * i], it.filt[2 * i + 1]));
    }
    const int qw = w >> 1, nq = qw * (h >> 1);
    for (int q = t; q < nq; q += kThreads) {
        const int qy = q / qw, qx = q - qy * qw;
        const int x = 2 * qx, y = 2 * qy;
        const uint16_t *c0 = &V[(y + E) * VW + x + VC];
        const uint16_t *pp[4] = {c0, c0 + 1, c0 + VW, c0 + VW + 1};
        int out[4];
#pragma unroll
        for (int p = 0; p < 4; p++) out[p] = VAT(pp[p], 0, 0);
        const bool lossless = (it.ll[y >> 2] >> (x >> 2)) & 1;
        if (!lossless) {
            if (kind != LR_KIND_NONE) {
                int s[4];
                if (kind == LR_KIND_SINGLE && pl == 0) {
#pragma unroll
                    for (int p = 0; p < 4; p++) s[p] = __mul24(out[p], fc);
                    sym_pairs<kNsYT, 8, VW>(pp, fs, s);
                } else if (kind == LR_KIND_SINGLE) {
                    const uint16_t *l0 = &Ld[(y + 2) * LS + x + 2];
                    const uint16_t *lq[4] = {l0, l0 + 1, l0 + LS, l0 + LS + 1};
#pragma unroll
                    for (int p = 0; p < 4; p++) s[p] = __mul24(out[p], fc) - __mul24((int)lq[p][0], fl);
                    sym_pairs<kNsUvT, 3, VW>(pp, fs, s);
                    tap_pairs<kNsUvYT, 6, LS>(lq, fs + 3, s);
                } else {
                    // multi-class NS-Wiener or PC-Wiener: the filter of the quad's 4x4 block
                    const int c = cls[(y >> 2) * 16 + (x >> 2)];
                    const uint4 *fp = (const uint4 *)(kind == LR_KIND_MULTI ? &sm.s_mfp[c * 8] : &sm.s_pcp[c * 8]);
                    const uint4 f0 = fp[0], f1 = fp[1];
                    const uint32_t fm[8] = {f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z, f1.w};
                    if (kind == LR_KIND_MULTI) {
                        const int fcm = sm.s_mF[c];
#pragma unroll
                        for (int p = 0; p < 4; p++) s[p] = __mul24(out[p], fcm);
                        sym_pairs<kNsYT, 8, VW>(pp, fm, s);
                    } else {
                        const int f12 = s16((int)fm[6]);  // centre tap
#pragma unroll
                        for (int p = 0; p < 4; p++) s[p] = __mul24(out[p], f12);
                        sym_pairs<kPcT, 6, VW>(pp, fm, s);
                    }
                }
#pragma unroll
                for (int p = 0; p < 4; p++) out[p] = iclip((s[p] + 64) >> 7, 0, bdmax);
            }
            if (gdf) {
                const uint16_t *cp = &cells[((y >> 1) * CS + (x >> 1)) * 4];
                int gs[4];
#pragma unroll
                for (int d = 0; d < 4; d++) gs[d] = cp[d] + cp[4 + d] + cp[CS * 4 + d] + cp[CS * 4 + 4 + d];
                const int c = (gs[0] <= gs[1]) | ((gs[2] <= gs[3]) << 1);
                int base[3] = {gbias[0], gbias[1], gbias[2]};
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    const int kk = d + 18;
                    const int v = dmin(gs[d] >> (4 - up), (int)sm.s_alpha[kk * 4 + c]);
#pragma unroll
                    for (int k = 0; k < 3; k++) base[k] += s16(v) * s16(sm.s_weight[k * 88 + kk * 4 + c]);
                }
                int iv[4][3];
#pragma unroll
                for (int p = 0; p < 4; p++)
#pragma unroll
                    for (int k = 0; k < 3; k++) iv[p][k] = base[k];
                // the 18 symmetric taps two at a time in packed 16-bit lanes
                // (|sample - m| << up <= 1023, alpha <= 511, weights in
                // [-512, 511]), with a 16-bit dot product per weight set
                const uint32_t *apk = &sm.apk[c * 9];
                const uint32_t *wpk = &sm.wpk[c * 27];
#pragma unroll
                for (int q2 = 0; q2 < 9; q2++) {
                    const s16x2 al = __builtin_bit_cast(s16x2, apk[q2]);
                    const s16x2 nal = -al;
                    const s16x2 w0 = __builtin_bit_cast(s16x2, wpk[q2]);
                    const s16x2 w1 = __builtin_bit_cast(s16x2, wpk[9 + q2]);
                    const s16x2 w2 = __builtin_bit_cast(s16x2, wpk[18 + q2]);
                    constexpr s16x2 lo = {-512, -512}, hi = {511, 511};
                    // stage by stage over the four pixels: consecutive packed
                    // operations are independent (a dependent pair would need a wait state)
                    s16x2 da[4], db[4], v[4];
#pragma unroll
                    for (int p = 0; p < 4; p++) {
                        const short m = (short)VAT(pp[p], 0, 0);
                        const s16x2 mm = {m, m};
                        const s16x2 A = {(short)VAT(pp[p], -kGdfT[2 * q2].dy, -kGdfT[2 * q2].dx),
                                         (short)VAT(pp[p], -kGdfT[2 * q2 + 1].dy, -kGdfT[2 * q2 + 1].dx)};
                        const s16x2 B = {(short)VAT(pp[p], kGdfT[2 * q2].dy, kGdfT[2 * q2].dx),
                                         (short)VAT(pp[p], kGdfT[2 * q2 + 1].dy, kGdfT[2 * q2 + 1].dx)};
                        if constexpr (down != 0) {
                            // samples (not differences) go to the 10-bit domain
                            da[p] = (A >> (short)down) - (mm >> (short)down);
                            db[p] = (B >> (short)down) - (mm >> (short)down);
                        } else {
                            da[p] = (A - mm) << (short)up;
                            db[p] = (B - mm) << (short)up;
                        }
                    }
#pragma unroll
                    for (int p = 0; p < 4; p++) da[p] = __builtin_elementwise_max(da[p], nal);
#pragma unroll
                    for (int p = 0; p < 4; p++) db[p] = __builtin_elementwise_max(db[p], nal);
#pragma unroll
                    for (int p = 0; p < 4; p++) da[p] = __builtin_elementwise_min(da[p], al);
#pragma unroll
                    for (int p = 0; p < 4; p++) db[p] = __builtin_elementwise_min(db[p], al);
#pragma unroll
                    for (int p = 0; p < 4; p++) v[p] = __builtin_elementwise_max(da[p] + db[p], lo);
#pragma unroll
                    for (int p = 0; p < 4; p++) v[p] = __builtin_elementwise_min(v[p], hi);
#pragma unroll
                    for (int p = 0; p < 4; p++) {
                        iv[p][0] = __builtin_amdgcn_sdot2(v[p], w0, iv[p][0], false);
                        iv[p][1] = __builtin_amdgcn_sdot2(v[p], w1, iv[p][1], false);
                        iv[p][2] = __builtin_amdgcn_sdot2(v[p], w2, iv[p][2], false);
                    }
                }
                int idx[4];
                if (intra)
                    gdf_index<8>(iv, idx);
                else
                    gdf_index<5>(iv, idx);
                const int8_t *err = sm.s_err + (intra ? kGdfErrBase8 : kGdfErrBase5);
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    // (|diff| + rnd) >> sh with diff's sign, as a floor (sh = 12 - bd >= 2)
                    constexpr int sh = 12 - bd;
                    const int diff = __mul24((int)err[idx[p]], a.gdf_scale);
                    if constexpr (sh == 0)
                        out[p] = iclip(out[p] + diff, 0, bdmax);  // 12 bit: no rounding shift
                    else
                        out[p] = iclip(out[p] + ((diff + (1 << (sh - 1)) + (diff >> 31)) >> sh), 0, bdmax);
                }
            }
        }
        P *d0 = (P *)(dbase + __umul24(y, stride)) + x;
        P *d1 = (P *)(dbase + __umul24(y + 1, stride)) + x;
        d0[0] = (P)out[0];
        d0[1] = (P)out[1];
        d1[0] = (P)out[2];
        d1[1] = (P)out[3];
    }
#undef VAT
}

// at least 5 waves per SIMD: the body is VALU/LDS-latency bound, so the
// register budget is capped (96 VGPRs) in exchange for occupancy (LDS allows
// 5 workgroups per CU)
template <typename P, int BD>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(5, 8))) void lr_kernel(const LrArgs a) {
    __shared__ LrSmem sm;
    lr_body<P, BD>(a, blockIdx.x, sm);
}

// Batched launch: every item (restoration units and plane copies) of every
// frame of one bit depth.
template <typename P, int BD, bool kLoop>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(5, 8))) void lr_batch(const LrArgs *__restrict__ args,
                                                                                      BatchSched s) {
    __shared__ LrSmem sm;
    for_each_item<kLoop>(s, [&](int f, int t) {
        lr_body<P, BD>(const_arg(args, f), t, sm);
        __syncthreads();
    });
}

// Mixed launch: the items of every bit-depth group's frames (8-, 10-, 12-bit bodies).
template <bool kLoop>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(5, 8))) void lr_mix(DevArgs3<LrArgs> args,
                                                                                    MixSched s) {
    __shared__ LrSmem sm;
    for_each_mixed_item<kLoop>(s, [&](int g, int f, int t) {
        if (g == 0)
            lr_body<uint8_t, 8>(const_arg(args.p[0], f), t, sm);
        else if (g == 1)
            lr_body<uint16_t, 10>(const_arg(args.p[1], f), t, sm);
        else
            lr_body<uint16_t, 12>(const_arg(args.p[2], f), t, sm);
        __syncthreads();
    });
}

void launch_lr_mix(DevArgs3<LrArgs> args, const MixSched &s, int grid, hipStream_t st) {
    if (s.first[3] <= 0) return;
    if (grid < s.first[3])
        R2G_LAUNCH((lr_mix<true>), dim3(grid), dim3(kThreads), 0, st, args, s);
    else
        R2G_LAUNCH((lr_mix<false>), dim3(grid), dim3(kThreads), 0, st, args, s);
}

void launch_lr_batch(const LrArgs *d_args, const BatchSched &sc, int grid, int bitdepth, hipStream_t s) {
    if (sc.total <= 0) return;
    const bool loop = grid < sc.total;
#define R2G_LR_BATCH(P, BD)                                                                      \
    do {                                                                                         \
        if (loop)                                                                                \
            R2G_LAUNCH((lr_batch<P, BD, true>), dim3(grid), dim3(kThreads), 0, s, d_args, sc);  \
        else                                                                                     \
            R2G_LAUNCH((lr_batch<P, BD, false>), dim3(grid), dim3(kThreads), 0, s, d_args, sc); \
    } while (0)
    if (bitdepth == 12)
        R2G_LR_BATCH(uint16_t, 12);
    else if (bitdepth == 10)
        R2G_LR_BATCH(uint16_t, 10);
    else
        R2G_LR_BATCH(uint8_t, 8);
#undef R2G_LR_BATCH
}

void launch_lr(const LrArgs &a, int n_items, int bitdepth, hipStream_t s) {
    if (n_items <= 0) return;
    if (bitdepth == 12)
        R2G_LAUNCH((lr_kernel<uint16_t, 12>), dim3(n_items), dim3(kThreads), 0, s, a);
    else if (bitdepth == 10)
        R2G_LAUNCH((lr_kernel<uint16_t, 10>), dim3(n_items), dim3(kThreads), 0, s, a);
    else
        R2G_LAUNCH((lr_kernel<uint8_t, 8>), dim3(n_items), dim3(kThreads), 0, s, a);
}

// Device copies of the tables, made once per device.
int lr_tables(int device, LrTables *out) {
    static LrTables cache[16];
    static bool ready[16] = {false};
    if (device < 0 || device >= 16) return RAV2D_EINVAL;
    if (!ready[device]) {
        // PC-Wiener filters as packed tap pairs for the 16-bit dot products
        static uint32_t pc_pairs[4][64][8];
        for (int wi = 0; wi < 4; wi++)
            for (int c = 0; c < 64; c++) {
                const int16_t *f = R2T_PC_WIENER_FILTERS[wi][c];
                for (int i = 0; i < 8; i++) {
                    const int lo = 2 * i < 13 ? f[2 * i] : 0, hi = 2 * i + 1 < 13 ? f[2 * i + 1] : 0;
                    pc_pairs[wi][c][i] = (uint32_t)(uint16_t)lo | ((uint32_t)(uint16_t)hi << 16);
                }
            }
        // the class of every feature index per sub-classification table
        static uint8_t cls_pc[4][4096], cls_ns[4][7][4096];
        for (int wi = 0; wi < 4; wi++)
            for (int li = 0; li < 4096; li++) {
                const int c = R2T_PC_WIENER_LUT_TO_CLASS[li];
                cls_pc[wi][li] = R2T_PC_WIENER_SUB_CLASSIFY[wi][c];
                for (int n = 0; n < 7; n++) cls_ns[wi][n][li] = R2T_PC_WIENER_SUB_CLASSIFY_NS[wi][n][c];
            }
        struct Src {
            const void *p;
            size_t n;
        } src[] = {
            {R2T_PC_WIENER_LUT_TO_CLASS, sizeof(R2T_PC_WIENER_LUT_TO_CLASS)},
            {R2T_PC_WIENER_FILTERS, sizeof(R2T_PC_WIENER_FILTERS)},
            {R2T_PC_WIENER_SUB_CLASSIFY, sizeof(R2T_PC_WIENER_SUB_CLASSIFY)},
            {R2T_PC_WIENER_SUB_CLASSIFY_NS, sizeof(R2T_PC_WIENER_SUB_CLASSIFY_NS)},
            {R2T_GDF_ALPHA, sizeof(R2T_GDF_ALPHA)},
            {R2T_GDF_WEIGHT, sizeof(R2T_GDF_WEIGHT)},
            {R2T_GDF_BIAS, sizeof(R2T_GDF_BIAS)},
            {R2T_GDF_INTRA_ERROR, sizeof(R2T_GDF_INTRA_ERROR)},
            {R2T_GDF_INTER_ERROR, sizeof(R2T_GDF_INTER_ERROR)},
            {pc_pairs, sizeof(pc_pairs)},
            {cls_pc, sizeof(cls_pc)},
            {cls_ns, sizeof(cls_ns)},
        };
        constexpr int kN = sizeof(src) / sizeof(src[0]);
        size_t total = 0, off[kN];
        for (int i = 0; i < kN; i++) {
            off[i] = total;
            total += (src[i].n + 255) & ~(size_t)255;
        }
        uint8_t *d = nullptr;
        if (hipMalloc((void **)&d, total) != hipSuccess) return RAV2D_ENOMEM;
        for (int i = 0; i < kN; i++)
            if (hipMemcpy(d + off[i], src[i].p, src[i].n, hipMemcpyHostToDevice) != hipSuccess) return RAV2D_EDEVICE;
        LrTables &t = cache[device];
        t.lut_to_class = d + off[0];
        t.pc_filters = (const int16_t *)(d + off[1]);
        t.pc_sub = d + off[2];
        t.ns_sub = d + off[3];
        t.gdf_alpha = (const uint16_t *)(d + off[4]);
        t.gdf_weight = (const int16_t *)(d + off[5]);
        t.gdf_bias = (const int16_t *)(d + off[6]);
        t.gdf_intra = (const int8_t *)(d + off[7]);
        t.gdf_inter = (const int8_t *)(d + off[8]);
        t.pc_pairs = (const uint32_t *)(d + off[9]);
        t.cls_pc = d + off[10];
        t.cls_ns = d + off[11];
        ready[device] = true;
    }
    *out = cache[device];
    return RAV2D_OK;
}

}  // namespace r2g
