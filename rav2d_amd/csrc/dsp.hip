// Per-family DSP entries: the function-table entries of rav2d's filter-stage
// DSP contexts (crates/rav2d/src/dsp.rs:9-14, 86-109) as single launches over
// one call's unit, for table-level drop-in and unit parity.  The frame-level
// entries (runtime.hip) are the fast path; these kernels favour a direct
// reading of each reference function: one lane per output sample (or per
// segment / block), the whole input read before anything is written (the
// reference works from line backups), side arrays passed by value.
//
// Film-grain template generation lives in grain.hip (it shares the set-up's
// wavefront AR filter); the entries here call it.
#include <string.h>

#include "deblock_core.h"
#include "frame.h"
#include "tables.h"

namespace r2g {

__device__ __forceinline__ int round2i(int x, int shift) { return (x + ((1 << shift) >> 1)) >> shift; }

template <typename P> __device__ __forceinline__ int sample_at(const uint8_t *base, int64_t stride, int y, int x) {
    return (int)((const P *)(base + (int64_t)y * stride))[x];
}
template <typename P> __device__ __forceinline__ void put_sample(uint8_t *base, int64_t stride, int y, int x, int v) {
    ((P *)(base + (int64_t)y * stride))[x] = (P)v;
}

// ---------------------------------------------------------------- deblock
struct DspDeblockArgs {
    uint8_t *dst;
    int64_t stride;
    uint16_t vmask[4];
    uint16_t ll[2];
    uint16_t q[16], side[16];
    int32_t dir, uv, edge, bdmax;
};

// deblock_{h,v}_sb64{y,uv}_bd (deblock.rs:303-551): lane i filters segment i
// (4 lines of up to 8 samples each side); segments touch disjoint lines.
template <typename P> __global__ __launch_bounds__(64) void dsp_deblock_sb(DspDeblockArgs a) {
    const int i = threadIdx.x;
    if (i >= 16) return;
    const uint32_t bit = 1u << i;
    const uint32_t vm = (uint32_t)a.vmask[0] | a.vmask[1] | a.vmask[2] | (a.uv ? 0u : (uint32_t)a.vmask[3]);
    if (!(vm & bit)) return;
    int idx;
    if (a.uv)
        idx = (a.vmask[2] & bit) ? 2 : ((a.vmask[1] & bit) != 0);
    else
        idx = (a.vmask[3] & bit) ? 3 : (a.vmask[2] & bit) ? 2 : ((a.vmask[1] & bit) != 0);
    const int mwp = byte_at(a.uv ? kMaxWidthUV : kMaxWidthY, idx);
    const int mwn = a.edge ? dmin(a.uv ? 2 : 6, mwp) : mwp;
    // step along the edge (sa) and across it (sb), in samples and rows
    const bool vert = a.dir == 0;
    auto at = [&](int line, int k) -> int {  // line of the segment, k - 8 across the edge
        const int along = 4 * i + line, across = k - 8;
        return vert ? sample_at<P>(a.dst, a.stride, along, across) : sample_at<P>(a.dst, a.stride, across, along);
    };
    int L[4][16];
#pragma unroll
    for (int l = 0; l < 4; l++)
#pragma unroll
        for (int k = 0; k < 16; k++) L[l][k] = at(l, k);
    const uint32_t q = a.q[i], side = a.side[i];
    const int width = filter_choice(L[0], L[3], mwn, mwp, q, side);
    const int wn = dmin(width, mwn), wp = width;
    if (wp < 1) return;
    const bool pos_ll = (a.ll[1] & bit) != 0, neg_ll = (a.ll[0] & bit) != 0;
    const int clampv = (int)q * byte_at(kQThreshMults, width - 1);
    const int mneg = byte_at(kWMult, wn - 1), mpos = byte_at(kWMult, wp - 1);
    for (int l = 0; l < 4; l++) {
        const int along = 4 * i + l;
        const int delta = iclip(4 * (3 * (L[l][8] - L[l][7]) - (L[l][9] - L[l][6])), -clampv, clampv);
        auto put = [&](int k, int v) {
            if (vert)
                put_sample<P>(a.dst, a.stride, along, k - 8, v);
            else
                put_sample<P>(a.dst, a.stride, k - 8, along, v);
        };
        if (!neg_ll)
            for (int j = 0; j < wn; j++)
                put(7 - j, iclip(L[l][7 - j] + ((delta * mneg * (wn - j) + 1024) >> 11), 0, a.bdmax));
        if (!pos_ll)
            for (int j = 0; j < wp; j++)
                put(8 + j, iclip(L[l][8 + j] - ((delta * mpos * (wp - j) + 1024) >> 11), 0, a.bdmax));
    }
}

// ---------------------------------------------------------------- CDEF
struct DspCdefDirArgs {
    const uint8_t *img;
    int64_t stride;
    int32_t bdm8;
    int32_t *dir_out;
    uint32_t *var_out;
};

// cdef_find_dir_bd (cdef.rs:134-216): lane d builds the partial sums of
// direction d from the staged block and its cost (wrapping u32 arithmetic);
// lane 0 picks the best.
template <typename P> __global__ __launch_bounds__(64) void dsp_cdef_dir(DspCdefDirArgs a) {
    __shared__ int px[8][8];
    __shared__ uint32_t cost[8];
    const int t = threadIdx.x;
    px[t >> 3][t & 7] = (sample_at<P>(a.img, a.stride, t >> 3, t & 7) >> a.bdm8) - 128;
    __syncthreads();
    const uint32_t div[7] = {840, 420, 280, 210, 168, 140, 120};
    auto sq = [](int v) { return (uint32_t)v * (uint32_t)v; };
    if (t < 8) {
        int ps[15];
        for (int k = 0; k < 15; k++) ps[k] = 0;
        // the line index of (y, x) for this direction (cdef.rs:145-157)
        for (int y = 0; y < 8; y++)
            for (int x = 0; x < 8; x++) {
                int k;
                switch (t) {
                    case 0: k = y + x; break;
                    case 1: k = y + (x >> 1); break;
                    case 2: k = y; break;
                    case 3: k = 3 + y - (x >> 1); break;
                    case 4: k = 7 + y - x; break;
                    case 5: k = 3 - (y >> 1) + x; break;
                    case 6: k = x; break;
                    default: k = (y >> 1) + x; break;
                }
                ps[k] += px[y][x];
            }
        uint32_t c = 0;
        if (t == 2 || t == 6) {
            for (int n = 0; n < 8; n++) c += sq(ps[n]);
            c *= 105;
        } else if (t == 0 || t == 4) {
            for (int n = 0; n < 7; n++) c += (sq(ps[n]) + sq(ps[14 - n])) * div[n];
            c += sq(ps[7]) * 105;
        } else {
            for (int m = 0; m < 5; m++) c += sq(ps[3 + m]);
            c *= 105;
            for (int m = 0; m < 3; m++) c += (sq(ps[m]) + sq(ps[10 - m])) * div[2 * m + 1];
        }
        cost[t] = c;
    }
    __syncthreads();
    if (t == 0) {
        int best = 0;
        uint32_t bc = cost[0];
        for (int n = 1; n < 8; n++)
            if (cost[n] > bc) {
                bc = cost[n];
                best = n;
            }
        *a.dir_out = best;
        *a.var_out = (bc - cost[best ^ 4]) >> 10;
    }
}

struct DspCdefArgs {
    uint8_t *dst;
    const uint8_t *left, *top, *bottom;
    int64_t stride;
    int32_t pri, sec, dir, damping, w, h, edges, bdm8;
    int8_t off[12][4];  // CDEF_DIRECTIONS as (dy0, dx0, dy1, dx1) per table row
};

constexpr int kNoSample = -32768;  // the reference's INT16_MIN padding

// cdef_padding (cdef.rs:37-114): the sample at (y, x) around the block, or
// INT16_MIN where the neighbour does not exist.
template <typename P> __device__ __forceinline__ int cdef_tap(const DspCdefArgs &a, int y, int x) {
    const bool L = a.edges & RAV2D_EDGE_LEFT, R = a.edges & RAV2D_EDGE_RIGHT;
    const bool T = a.edges & RAV2D_EDGE_TOP, B = a.edges & RAV2D_EDGE_BOTTOM;
    if ((y < 0 && !T) || (y >= a.h && !B) || (x < 0 && !L) || (x >= a.w && !R)) return kNoSample;
    if (y < 0) return sample_at<P>(a.top, a.stride, y + 2, x);
    if (y >= a.h) return sample_at<P>(a.bottom, a.stride, y - a.h, x);
    if (x < 0) return (int)((const P *)a.left)[y * 2 + 2 + x];
    return sample_at<P>(a.dst, a.stride, y, x);
}

__device__ __forceinline__ int cdef_constrain(int diff, int threshold, int shift) {
    const int ad = abs(diff);
    return apply_sign(dmin(ad, dmax(0, threshold - (ad >> shift))), diff);
}

// cdef_filter_block (cdef.rs:320-460), one lane per sample: primary taps
// along direction dir (table row dir + 2), secondary taps along dir + 2 and
// dir - 2 (rows dir + 4 and dir); the result is clipped to the taps' range
// when both strengths are on.
template <typename P> __global__ __launch_bounds__(64) void dsp_cdef_filter(DspCdefArgs a) {
    const int t = threadIdx.x, y = t / a.w, x = t - y * a.w;
    const bool on = y < a.h;
    int out = 0;
    if (on) {
        const int px = cdef_tap<P>(a, y, x);
        int sum = 0, mx = px, mn = px;
        const bool pri = a.pri != 0, sec = a.sec != 0;
        const int pri_shift = pri ? dmax(0, a.damping - ulog2((uint32_t)a.pri)) : 0;
        const int sec_shift = sec ? a.damping - ulog2((uint32_t)a.sec) : 0;
        int tk = 4 - ((a.pri >> a.bdm8) & 1);
        for (int k = 0; k < 2; k++) {
            if (pri) {
                const int dy = a.off[a.dir + 2][0 + 2 * k], dx = a.off[a.dir + 2][1 + 2 * k];
                const int p0 = cdef_tap<P>(a, y + dy, x + dx), p1 = cdef_tap<P>(a, y - dy, x - dx);
                sum += tk * cdef_constrain(p0 - px, a.pri, pri_shift);
                sum += tk * cdef_constrain(p1 - px, a.pri, pri_shift);
                tk = (tk & 3) | 2;
                mn = dmin(dmin(p0, p1), mn);
                mx = dmax(dmax(p0, p1), mx);
            }
            if (sec) {
                const int st = 2 - k;
                for (int e = 0; e < 2; e++) {
                    const int d = e ? a.dir : a.dir + 4;
                    const int dy = a.off[d][0 + 2 * k], dx = a.off[d][1 + 2 * k];
                    const int s0 = cdef_tap<P>(a, y + dy, x + dx), s1 = cdef_tap<P>(a, y - dy, x - dx);
                    sum += st * cdef_constrain(s0 - px, a.sec, sec_shift);
                    sum += st * cdef_constrain(s1 - px, a.sec, sec_shift);
                    mn = dmin(dmin(s0, s1), mn);
                    mx = dmax(dmax(s0, s1), mx);
                }
            }
        }
        out = px + ((sum - (sum < 0) + 8) >> 4);
        if (pri && sec) out = iclip(out, mn, mx);
    }
    __syncthreads();  // every lane has read its taps before any writes
    if (on) put_sample<P>(a.dst, a.stride, y, x, out);
}

// ---------------------------------------------------------------- CCSO
struct DspCcsoPrepArgs {
    uint8_t *dst;
    const uint8_t *src, *left, *top, *bottom;
    int64_t dst_stride, src_stride;
    int32_t shift, dy, dx, quant_step, bo_only, edges, w, h, ss_hor, ss_ver;
    uint32_t edge_clf;
};

// ccso_padding (ccso.rs:26-64): the luma sample at (y, x) of the block's
// neighbourhood, coordinates clamped to the neighbours that exist.
template <typename P> __device__ __forceinline__ int ccso_luma(const DspCcsoPrepArgs &a, int y, int x) {
    const int pw = a.w << a.ss_hor, ph = a.h << a.ss_ver;
    const int x_min = (a.edges & RAV2D_EDGE_LEFT) ? -2 : 0, x_max = pw - 1 + ((a.edges & RAV2D_EDGE_RIGHT) ? 2 : 0);
    const int y_min = (a.edges & RAV2D_EDGE_TOP) ? -2 : 0, y_max = ph - 1 + ((a.edges & RAV2D_EDGE_BOTTOM) ? 2 : 0);
    const int sy = iclip(y, y_min, y_max), sx = iclip(x, x_min, x_max);
    if (sy < 0) return sample_at<P>(a.top, a.src_stride, 2 + sy, sx);
    if (sy >= ph) return sample_at<P>(a.bottom, a.src_stride, sy - ph, sx);
    if (sx < 0) return (int)((const P *)a.left)[sy * 2 + 2 + sx];
    return sample_at<P>(a.src, a.src_stride, sy, sx);
}

__device__ __forceinline__ uint32_t ccso_score(int diff, int quant_step, uint32_t edge_clf) {
    if (diff > quant_step && edge_clf == 0) return 2;
    if (diff < -quant_step) return 0;
    return 1;
}

// ccso_prep (ccso.rs:115-186): one lane per sample, looping over the block.
template <typename P> __global__ __launch_bounds__(256) void dsp_ccso_prep(DspCcsoPrepArgs a) {
    for (int i = threadIdx.x; i < a.w * a.h; i += blockDim.x) {
        const int y = i / a.w, x = i - y * a.w;
        const int ly = y << a.ss_ver, lx = x << a.ss_hor;
        const int c = ccso_luma<P>(a, ly, lx);
        const uint8_t band = (uint8_t)((uint32_t)c >> a.shift);
        uint8_t v = band;
        if (!a.bo_only) {
            const uint32_t c0 = ccso_score(ccso_luma<P>(a, ly + a.dy, lx + a.dx) - c, a.quant_step, a.edge_clf);
            const uint32_t c1 = ccso_score(ccso_luma<P>(a, ly - a.dy, lx - a.dx) - c, a.quant_step, a.edge_clf);
            v = (uint8_t)((c0 << 5) | (c1 << 3)) | band;
        }
        a.dst[(int64_t)y * a.dst_stride + x] = v;
    }
}

struct DspCcsoAddArgs {
    uint8_t *dst;
    const uint8_t *idx;
    int64_t dst_stride, idx_stride;
    uint8_t offs[64];
    int8_t lut[8];
    uint16_t ll[16];
    int32_t w, h, bdmax;
};

// ccso_add (ccso.rs:215-254)
template <typename P> __global__ __launch_bounds__(256) void dsp_ccso_add(DspCcsoAddArgs a) {
    for (int i = threadIdx.x; i < a.w * a.h; i += blockDim.x) {
        const int y = i / a.w, x = i - y * a.w;
        if ((a.ll[y >> 2] >> (x >> 2)) & 1) continue;
        const uint8_t c = a.idx[(int64_t)y * a.idx_stride + x];
        const int bi = c >> 1, hi = c & 1;
        const uint8_t packed = bi < 64 ? a.offs[bi] : 0;
        const int v = sample_at<P>(a.dst, a.dst_stride, y, x) + a.lut[7 & (packed >> (4 * hi))];
        put_sample<P>(a.dst, a.dst_stride, y, x, iclip(v, 0, a.bdmax));
    }
}

// ---------------------------------------------------------------- LR
// The stripe's "virtual image" the reference's row backups build
// (backup_row / backup_row_lpf, looprestoration.rs:54-109, and the row walks
// of the stripe kernels): rows inside the stripe from p (left of column 0 the
// caller's left buffer), the rows above / below from top / bottom (rows -1 /
// h are their second / first row, further rows repeat them), without the
// neighbour the first / last row and column repeat.
struct Stripe {
    const uint8_t *p, *left, *top, *bottom;
    int64_t stride;
    int32_t w, h, edges;
};

template <typename P> __device__ __forceinline__ int vimg(const Stripe &s, int y, int x) {
    const bool T = s.edges & RAV2D_EDGE_TOP, B = s.edges & RAV2D_EDGE_BOTTOM;
    const bool L = s.edges & RAV2D_EDGE_LEFT, R = s.edges & RAV2D_EDGE_RIGHT;
    const uint8_t *row;
    bool inside = false;
    if (y < 0) {
        if (T) {
            row = s.top + (y == -1 ? s.stride : 0);
        } else {
            row = s.p;
            inside = true;
            y = 0;
        }
    } else if (y >= s.h) {
        if (B) {
            row = s.bottom + (y == s.h ? 0 : s.stride);
        } else {
            y = s.h - 1;
            row = s.p + (int64_t)y * s.stride;
            inside = true;
        }
    } else {
        row = s.p + (int64_t)y * s.stride;
        inside = true;
    }
    if (x < 0) {
        if (!L) return ((const P *)row)[0];
        if (inside) return ((const P *)s.left)[y * 6 + 6 + x];
        return ((const P *)row)[x];
    }
    if (x >= s.w && !R) return ((const P *)row)[s.w - 1];
    return ((const P *)row)[x];
}

// The collocated luma of the chroma stripe (backup_row_luma, :408-522, 4:2:0):
// chroma row y averages luma rows 2y, 2y + 1 (cfl_ds 0: with the next column,
// 1: vertically, 2: the top-left sample); above the stripe row -1 takes the
// two luma rows above it (luma_top), row -2 the first of them twice; below,
// row h the two rows of luma_bottom, row h + 1 its second twice; columns out
// of the unit repeat the first / last pair when the neighbour is missing.
struct LumaRows {
    const uint8_t *luma, *top, *bottom;
    int64_t stride;
    int32_t ds;
};

template <typename P> __device__ __forceinline__ int vluma(const Stripe &s, const LumaRows &l, int y, int x) {
    const bool T = s.edges & RAV2D_EDGE_TOP, B = s.edges & RAV2D_EDGE_BOTTOM;
    const bool Lft = s.edges & RAV2D_EDGE_LEFT, R = s.edges & RAV2D_EDGE_RIGHT;
    const uint8_t *r0, *r1;
    if (y < 0 && T) {
        r0 = l.top;
        r1 = y == -1 ? l.top + l.stride : l.top;
    } else if (y >= s.h && B) {
        r0 = y == s.h ? l.bottom : l.bottom + l.stride;
        r1 = l.bottom + l.stride;
    } else {
        const int yy = iclip(y, 0, s.h - 1);
        r0 = l.luma + (int64_t)(2 * yy) * l.stride;
        r1 = r0 + l.stride;
    }
    const int lw = 2 * s.w;
    int lx = 2 * x;
    if (lx < 0 && !Lft) lx = 0;
    if (lx >= lw && !R) lx = lw - 2;
    const P *a = (const P *)r0, *b = (const P *)r1;
    if (l.ds == 0) return ((int)a[lx] + b[lx] + a[lx + 1] + b[lx + 1]) >> 2;
    if (l.ds == 1) return ((int)a[lx] + b[lx]) >> 1;
    return a[lx];
}

constexpr int8_t kNsY[16][2] = {{1, 0}, {0, 1}, {2, 0}, {0, 2}, {1, 1}, {-1, 1}, {2, 1}, {2, -1},
                                 {1, 2}, {1, -2}, {3, 0}, {0, 3}, {4, 0}, {0, 4}, {3, 3}, {3, -3}};
constexpr int8_t kNsUv[6][2] = {{1, 0}, {0, 1}, {1, 1}, {-1, 1}, {2, 0}, {0, 2}};
constexpr int8_t kNsUvFromY[12][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, -1},
                                       {-1, 1}, {1, -1}, {2, 0}, {-2, 0}, {0, 2}, {0, -2}};
constexpr int8_t kPc[12][2] = {{1, 0}, {0, 1}, {2, 0}, {0, 2}, {1, 1}, {-1, 1},
                                {2, 1}, {2, -1}, {1, 2}, {1, -2}, {3, 0}, {0, 3}};
constexpr int8_t kGdfCoords[18][2] = {{6, 0}, {5, 0}, {4, 0}, {3, 0}, {2, 1}, {2, 0}, {2, -1}, {1, 2}, {1, 1},
                                       {1, 0}, {1, -1}, {1, -2}, {0, 6}, {0, 5}, {0, 4}, {0, 3}, {0, 2}, {0, 1}};

// get_qval_given_tskip (looprestoration.rs:40-49)
__device__ __forceinline__ int qval_given_tskip(int qstep, int tskip, int i, int bdm8) {
    constexpr int16_t kW[4][3] = {{-527, 15325, 321}, {26436, -17705, 17905}, {366, -147, -194}, {202, -267, -179}};
    constexpr int16_t kO[4] = {-547, -21565, -573, -680};
    qstep = (qstep + ((1 << bdm8) >> 1)) >> bdm8;
    const int prod = (tskip * qstep + 128) >> 8;
    const int q = kW[i][0] * (tskip << 5) + kW[i][1] * qstep + kW[i][2] * prod;
    return 255 * (kO[i] + apply_sign((abs(q) + (1 << 12)) >> 13, q));
}

struct DspLrArgs {
    Stripe s;
    LumaRows l;
    uint8_t *dst;  // == s.p (written after every read)
    int32_t kind, bdmax, bdm8, base_q;
    int8_t filt[18];
    uint16_t ll[17];
    uint16_t noskip[17];
    uint8_t sub[256];
    const uint8_t *lut_to_class;  // PC_WIENER_LUT_TO_CLASS (device)
    union {
        int8_t user[16][18];
        int16_t pre[64][13];
    } f;
};

// get_class_lut_idx (looprestoration.rs:195-271) of 4x4 block (bx, by).
template <typename P> __device__ int lr_class(const DspLrArgs &a, int bx, int by) {
    int f[3] = {0, 0, 0}, sk = 0;
    const int bh = a.s.h >> 2;
    for (int dy = -1; dy <= 4; dy++)
        for (int dx = -1; dx <= 4; dx++) {
            const int x = bx * 4 + dx, y = by * 4 + dy;
            const int m = vimg<P>(a.s, y, x);
            f[0] += abs(vimg<P>(a.s, y - 1, x) - 2 * m + vimg<P>(a.s, y + 1, x));
            f[1] += abs(vimg<P>(a.s, y - 1, x + 1) - 2 * m + vimg<P>(a.s, y + 1, x - 1));
            f[2] += abs(vimg<P>(a.s, y - 1, x - 1) - 2 * m + vimg<P>(a.s, y + 1, x + 1));
        }
    constexpr uint8_t kPixels[3] = {16, 4, 1};
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            const int e = (dy != 0) + (dx != 0);
            const int fx = iclip((bx & 15) + dx, 0, 15), fy = iclip(by + dy, 0, bh - 1);
            sk += kPixels[e] * (((a.noskip[fy] >> fx) & 1) == 0);
        }
    constexpr uint16_t kNorm[4] = {3739, 3273, 3074, 7};
    const int rnd = (1 << a.bdm8) >> 1;
    for (int i = 0; i < 3; i++) f[i] = (int)((uint32_t)f[i] * kNorm[i] + (uint32_t)rnd) >> a.bdm8;
    sk *= kNorm[3];
    int q = (dmax(0, qval_given_tskip(a.base_q, sk, 0, a.bdm8)) + (1 << 13)) >> 14;
    int li = (dmin(q, 255) >> 5) << 9;
    for (int i = 0; i < 3; i++) {
        q = (dmax(0, f[i] + qval_given_tskip(a.base_q, sk, i + 1, a.bdm8)) + (1 << 13)) >> 14;
        li |= (dmin(q, 255) >> 5) << (3 * (2 - i));
    }
    return a.sub[a.lut_to_class[li]];
}

// WienerFilterFn kinds (ns_wiener_single luma / chroma, ns_wiener_multi,
// pc_wiener), one lane per sample; classes first (one lane per 4x4 block).
template <typename P> __global__ __launch_bounds__(256) void dsp_lr_wiener(DspLrArgs a) {
    __shared__ uint8_t cls[256];
    __shared__ uint16_t out[64 * 64];
    const int w = a.s.w, h = a.s.h, t = threadIdx.x;
    if (a.kind >= 2) {
        for (int b = t; b < (w >> 2) * (h >> 2); b += blockDim.x) {
            const int by = b / (w >> 2), bx = b - by * (w >> 2);
            cls[b] = (uint8_t)lr_class<P>(a, bx, by);
        }
        __syncthreads();
    }
    for (int i = t; i < w * h; i += blockDim.x) {
        const int y = i / w, x = i - y * w;
        const int m = vimg<P>(a.s, y, x);
        int sum;
        if (a.kind == 0 || a.kind == 2) {
            const int8_t *fl = a.kind == 0 ? a.filt : a.f.user[cls[(y >> 2) * (w >> 2) + (x >> 2)]];
            sum = m << 7;
            for (int k = 0; k < 16; k++) {
                const int dy = kNsY[k][0], dx = kNsY[k][1];
                sum += (vimg<P>(a.s, y + dy, x + dx) + vimg<P>(a.s, y - dy, x - dx) - 2 * m) * fl[k];
            }
        } else if (a.kind == 3) {
            const int16_t *fl = a.f.pre[cls[(y >> 2) * (w >> 2) + (x >> 2)]];
            sum = m * fl[12];
            for (int k = 0; k < 12; k++) {
                const int dy = kPc[k][0], dx = kPc[k][1];
                sum += (vimg<P>(a.s, y + dy, x + dx) + vimg<P>(a.s, y - dy, x - dx)) * fl[k];
            }
        } else {
            sum = m << 7;
            for (int k = 0; k < 6; k++) {
                const int dy = kNsUv[k][0], dx = kNsUv[k][1];
                sum += (vimg<P>(a.s, y + dy, x + dx) + vimg<P>(a.s, y - dy, x - dx) - 2 * m) * a.filt[k];
            }
            const int lc = vluma<P>(a.s, a.l, y, x);
            for (int k = 0; k < 12; k++)
                sum += (vluma<P>(a.s, a.l, y + kNsUvFromY[k][0], x + kNsUvFromY[k][1]) - lc) * a.filt[6 + k];
        }
        const bool skip = (a.ll[y >> 2] >> (x >> 2)) & 1;
        out[i] = (uint16_t)(skip ? m : iclip((sum + 64) >> 7, 0, a.bdmax));
    }
    __syncthreads();  // the whole stripe has been read
    for (int i = t; i < w * h; i += blockDim.x) {
        const int y = i / w, x = i - y * w;
        put_sample<P>(a.dst, a.s.stride, y, x, out[i]);
    }
}

struct DspGdfArgs {
    Stripe s;
    int8_t *dst;
    int64_t dst_stride;
    int32_t up, down, alpha_base, weight_base, bias_idx, intra, err_base, scale;
    const uint16_t *alpha;
    const int16_t *weight, *bias;
    const int8_t *err_intra, *err_inter;
};

// gdf_prep (looprestoration.rs:1432-1643): the class of each 2x2 output pair
// from 4-direction gradients over rows y0 - 1 .. y0 + 2 and columns x0 - 1 ..
// x0 + 2, then 18 clipped symmetric features and 4 gradient features weighted
// into 3 sub-indices of the error table.
template <typename P> __global__ __launch_bounds__(256) void dsp_gdf_prep(DspGdfArgs a) {
    const int w = a.s.w, h = a.s.h;
    constexpr int kOff[4][2] = {{1, 0}, {0, 1}, {1, 1}, {-1, 1}};
    // samples in the 10-bit domain of the error tables (12-bit ones shifted down)
    auto V = [&](int r, int c) { return vimg<P>(a.s, r, c) >> a.down; };
    for (int i = threadIdx.x; i < w * h; i += blockDim.x) {
        const int y = i / w, x = i - y * w;
        const int y0 = y & ~1, x0 = x & ~1;
        int gs[4];
        for (int d = 0; d < 4; d++) {
            int g = 0;
            for (int r = y0 - 1; r <= y0 + 2; r++)
                for (int c = x0 - 1; c <= x0 + 2; c++) {
                    const int dy = kOff[d][0], dx = kOff[d][1];
                    g += abs(2 * V(r, c) - V(r - dy, c - dx) - V(r + dy, c + dx));
                }
            gs[d] = g;
        }
        const int cl = (gs[0] <= gs[1]) | ((gs[2] <= gs[3]) << 1);
        int iv[3];
        for (int k = 0; k < 3; k++) iv[k] = a.bias[a.bias_idx * 3 + k];
        for (int d = 0; d < 4; d++) {
            const int k = d + 18;
            const int v = dmin(gs[d] >> (4 - a.up), (int)a.alpha[a.alpha_base + k * 4 + cl]);
            for (int j = 0; j < 3; j++) iv[j] += v * a.weight[a.weight_base + j * 88 + k * 4 + cl];
        }
        const int m = V(y, x);
        for (int k = 0; k < 18; k++) {
            const int al = a.alpha[a.alpha_base + k * 4 + cl];
            const int dy = kGdfCoords[k][0], dx = kGdfCoords[k][1];
            const int above = iclip((V(y - dy, x - dx) - m) * (1 << a.up), -al, al);
            const int below = iclip((V(y + dy, x + dx) - m) * (1 << a.up), -al, al);
            const int v = iclip(above + below, -512, 511);
            for (int j = 0; j < 3; j++) iv[j] += v * a.weight[a.weight_base + j * 88 + k * 4 + cl];
        }
        int full = 0;
        for (int j = 0; j < 3; j++) {
            const int sv = iv[j] * a.scale;
            const int v = apply_sign((abs(sv) + (1 << 14)) >> 15, sv);
            full = full * (a.scale * 2) + iclip(v, -a.scale, a.scale - 1) + a.scale;
        }
        a.dst[(int64_t)y * a.dst_stride + x] = a.intra ? a.err_intra[a.err_base + full] : a.err_inter[a.err_base + full];
    }
}

struct DspGdfAddArgs {
    uint8_t *p;
    const int8_t *err;
    int64_t stride, err_stride;
    uint16_t ll[17];
    int32_t w, h, scale, shift, bdmax;
};

// gdf_add (looprestoration.rs:273-329)
template <typename P> __global__ __launch_bounds__(256) void dsp_gdf_add(DspGdfAddArgs a) {
    for (int i = threadIdx.x; i < a.w * a.h; i += blockDim.x) {
        const int y = i / a.w, x = i - y * a.w;
        if ((a.ll[y >> 2] >> (x >> 2)) & 1) continue;
        const int diff = a.err[(int64_t)y * a.err_stride + x] * a.scale;
        const int v = sample_at<P>(a.p, a.stride, y, x) + apply_sign((abs(diff) + ((1 << a.shift) >> 1)) >> a.shift, diff);
        put_sample<P>(a.p, a.stride, y, x, iclip(v, 0, a.bdmax));
    }
}

// ---------------------------------------------------------------- film grain
struct DspGrainArgs {
    uint8_t *dst;
    const uint8_t *src, *luma;
    int64_t stride, luma_stride;
    const uint8_t *scaling;
    const int16_t *lut;  // rows of 82
    Rav2dFilmGrainData fg;
    uint32_t seed;
    int32_t pw, bh, row_num, bitdepth, chroma, uv, is_id, sx, sy;
};

__device__ __forceinline__ uint32_t lfsr_draw(uint32_t bits, uint32_t &s) {
    const uint32_t bit = (s ^ (s >> 1) ^ (s >> 3) ^ (s >> 12)) & 1;
    s = (s >> 1) | (bit << 15);
    return (s >> (16 - bits)) & ((1u << bits) - 1);
}

// fgy_32x32xn / fguv_32x32xn (filmgrain.rs:248-592) over one row band: each
// workgroup takes 64 columns; two lanes walk the band's and the previous
// band's LFSR streams up to the workgroup's blocks (per block, per offset a
// 9-bit draw scaled by 3 - block_size then three 16-bit draws), then every
// lane computes its samples: the template window of its block, blended with
// the left / upper neighbours' windows in the overlap columns / rows.
template <typename P> __global__ __launch_bounds__(256) void dsp_grain_band(DspGrainArgs a) {
    __shared__ uint8_t offs[2][12][2];  // [band: 0 this, 1 previous][block - first][y, x]
    const Rav2dFilmGrainData &d = a.fg;
    // block width / height in this plane
    const int bs = 16 << d.block_size, pbs = bs >> a.sx, pbsy = bs >> a.sy;
    const int x0 = blockIdx.x * 64, x1 = dmin(x0 + 64, a.pw);
    const int b_first = dmax(x0 / pbs - 1, 0), b_last = (x1 - 1) / pbs;
    const int rows = 1 + (d.overlap_flag && a.row_num > 0);
    if (threadIdx.x < (unsigned)rows) {
        const int i = threadIdx.x, band = a.row_num - i;
        uint32_t s = a.seed;
        s ^= (uint32_t)(((band * 37 + 178) & 0xFF)) << 8;
        s ^= (uint32_t)((band * 173 + 105) & 0xFF);
        for (int b = 0; b <= b_last; b++)
            for (int n = 0; n < 2; n++) {
                const uint32_t v = ((uint32_t)(3 - d.block_size) * lfsr_draw(9, s)) >> 6;
                for (int k = 0; k < 3; k++) lfsr_draw(16, s);
                if (b >= b_first) offs[i][b - b_first][n] = (uint8_t)v;
            }
    }
    __syncthreads();
    const int bdm8 = a.bitdepth - 8, bdmax = (1 << a.bitdepth) - 1;
    const int gmin = -(128 << bdm8), gmax = (128 << bdm8) - 1;
    int minv = 0, maxv = bdmax;
    if (d.clip_to_restricted_range) {
        minv = 16 << bdm8;
        maxv = ((!a.chroma || a.is_id) ? 235 : 240) << bdm8;
    }
    // sample_lut (filmgrain.rs:231-246) for block (bx 0 this / 1 left, by 0 this / 1 upper)
    auto lut_at = [&](int b, int bx, int by, int lx, int ly) -> int {
        const uint8_t *o = offs[by][b - bx - b_first];
        const int offx = 3 + (2 >> a.sx) * (3 + o[1]), offy = 3 + (2 >> a.sy) * (3 + o[0]);
        return a.lut[(offy + ly + pbsy * by) * 82 + offx + lx + pbs * bx];
    };
    auto blend = [&](int old, int cur, int w0, int w1) { return iclip(round2i(old * w0 + cur * w1, 5), gmin, gmax); };
    const int wts[2][2][2] = {{{27, 17}, {17, 27}}, {{23, 22}, {0, 0}}};
    const int (*wx)[2] = wts[a.sx], (*wy)[2] = wts[a.sy];
    const int n = (x1 - x0) * a.bh;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int y = i / (x1 - x0), x = x0 + (i - y * (x1 - x0));
        const int b = x / pbs, lx = x - b * pbs;
        const int bw = dmin(pbs, a.pw - b * pbs);
        const int xstart = (d.overlap_flag && b > 0) ? dmin(2 >> a.sx, bw) : 0;
        const int ystart = (d.overlap_flag && a.row_num > 0) ? dmin(2 >> a.sy, a.bh) : 0;
        int g = lut_at(b, 0, 0, lx, y);
        if (lx < xstart) g = blend(lut_at(b, 1, 0, lx, y), g, wx[lx][0], wx[lx][1]);
        if (y < ystart) {
            int top = lut_at(b, 0, 1, lx, y);
            if (lx < xstart) top = blend(lut_at(b, 1, 1, lx, y), top, wx[lx][0], wx[lx][1]);
            g = blend(top, g, wy[y][0], wy[y][1]);
        }
        const int sv = sample_at<P>(a.src, a.stride, y, x);
        int val = sv;
        if (a.chroma) {
            const int lxx = x << a.sx, lyy = y << a.sy;
            const int l0 = sample_at<P>(a.luma, a.luma_stride, lyy, lxx);
            const int avg = a.sx ? (l0 + sample_at<P>(a.luma, a.luma_stride, lyy, lxx + 1) + 1) >> 1 : l0;
            if (!d.chroma_scaling_from_luma) {
                const int comb = avg * d.uv_luma_mult[a.uv] + sv * d.uv_mult[a.uv];
                val = iclip((comb >> 6) + d.uv_offset[a.uv] * (1 << bdm8), 0, bdmax);
            } else {
                val = avg;
            }
        }
        const int noise = round2i(a.scaling[val] * g, d.scaling_shift);
        put_sample<P>(a.dst, a.stride, y, x, iclip(sv + noise, minv, maxv));
    }
}

}  // namespace r2g

using namespace r2g;

namespace {

bool bad_depth(int bitdepth) {
    if (bitdepth != 8 && bitdepth != 10 && bitdepth != 12) {
        set_last_error("bit depth must be 8, 10 or 12");
        return true;
    }
    return false;
}

int bad(const char *msg) {
    set_last_error(msg);
    return RAV2D_EINVAL;
}

#define R2G_DSP_LAUNCH(K, G, B, S, A)                                   \
    do {                                                                \
        if (bitdepth > 8)                                               \
            hipLaunchKernelGGL((K<uint16_t>), dim3(G), dim3(B), 0, S, A); \
        else                                                            \
            hipLaunchKernelGGL((K<uint8_t>), dim3(G), dim3(B), 0, S, A);  \
        HIP_TRY(hipGetLastError());                                     \
    } while (0)

Stripe stripe_of(const void *p, ptrdiff_t stride, const void *left, const void *top, const void *bottom, int w, int h,
                 int edges) {
    Stripe s;
    s.p = (const uint8_t *)p;
    s.left = (const uint8_t *)left;
    s.top = (const uint8_t *)top;
    s.bottom = (const uint8_t *)bottom;
    s.stride = stride;
    s.w = w;
    s.h = h;
    s.edges = edges;
    return s;
}

}  // namespace

extern "C" {

int rav2d_gpu_deblock_sb(Rav2dGpuCtx *ctx, int dir, int uv, void *dst, ptrdiff_t stride, const uint16_t *vmask,
                         const uint16_t ll_mask[2], const uint16_t q_thr[16], const uint16_t side_thr[16], int edge,
                         int bitdepth, void *stream) {
    if (!ctx || !dst || !vmask || !ll_mask || !q_thr || !side_thr || (dir != 0 && dir != 1) || (uv != 0 && uv != 1))
        return bad("deblock_sb: invalid arguments");
    if (bad_depth(bitdepth)) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    DspDeblockArgs a;
    memset(&a, 0, sizeof(a));
    a.dst = (uint8_t *)dst;
    a.stride = stride;
    for (int i = 0; i < (uv ? 3 : 4); i++) a.vmask[i] = vmask[i];
    a.ll[0] = ll_mask[0];
    a.ll[1] = ll_mask[1];
    memcpy(a.q, q_thr, sizeof(a.q));
    memcpy(a.side, side_thr, sizeof(a.side));
    a.dir = dir;
    a.uv = uv;
    a.edge = edge != 0;
    a.bdmax = (1 << bitdepth) - 1;
    R2G_DSP_LAUNCH(dsp_deblock_sb, 1, 64, pick_stream(ctx, stream), a);
    return RAV2D_OK;
}

int rav2d_gpu_cdef_dir(Rav2dGpuCtx *ctx, const void *img, ptrdiff_t stride, int bitdepth, int32_t *dir_out,
                       uint32_t *var_out, void *stream) {
    if (!ctx || !img || !dir_out || !var_out) return bad("cdef_dir: invalid arguments");
    if (bad_depth(bitdepth)) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    DspCdefDirArgs a;
    a.img = (const uint8_t *)img;
    a.stride = stride;
    a.bdm8 = bitdepth - 8;
    a.dir_out = dir_out;
    a.var_out = var_out;
    R2G_DSP_LAUNCH(dsp_cdef_dir, 1, 64, pick_stream(ctx, stream), a);
    return RAV2D_OK;
}

int rav2d_gpu_cdef_filter(Rav2dGpuCtx *ctx, void *dst, ptrdiff_t stride, const void *left, const void *top,
                          const void *bottom, int pri_strength, int sec_strength, int dir, int damping, int w, int h,
                          int edges, int bitdepth, void *stream) {
    if (!ctx || !dst || !left || !top || !bottom || (w != 4 && w != 8) || (h != 4 && h != 8) || dir < 0 || dir > 7 ||
        pri_strength < 0 || sec_strength < 0 || (pri_strength == 0 && sec_strength == 0))
        return bad("cdef_filter: invalid arguments (w, h in {4, 8}, dir 0..7, a strength on)");
    if (bad_depth(bitdepth)) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    DspCdefArgs a;
    memset(&a, 0, sizeof(a));
    a.dst = (uint8_t *)dst;
    a.left = (const uint8_t *)left;
    a.top = (const uint8_t *)top;
    a.bottom = (const uint8_t *)bottom;
    a.stride = stride;
    a.pri = pri_strength;
    a.sec = sec_strength;
    a.dir = dir;
    a.damping = damping;
    a.w = w;
    a.h = h;
    a.edges = edges;
    a.bdm8 = bitdepth - 8;
    // CDEF_DIRECTIONS holds dy * 12 + dx (the reference's 12-wide padded block)
    for (int r = 0; r < 12; r++)
        for (int k = 0; k < 2; k++) {
            const int v = R2T_CDEF_DIRECTIONS[r][k];
            const int dy = (v + 12 * 3 + 6) / 12 - 3;
            a.off[r][2 * k] = (int8_t)dy;
            a.off[r][2 * k + 1] = (int8_t)(v - 12 * dy);
        }
    R2G_DSP_LAUNCH(dsp_cdef_filter, 1, 64, pick_stream(ctx, stream), a);
    return RAV2D_OK;
}

int rav2d_gpu_ccso_prep(Rav2dGpuCtx *ctx, uint8_t *dst, ptrdiff_t dst_stride, const void *src, ptrdiff_t src_stride,
                        const void *left, const void *top, const void *bottom, int max_band_log2, int ext_filter,
                        int quant_step, uint32_t edge_clf, int bo_only, int edges, int w, int h, int ss_hor,
                        int ss_ver, int bitdepth, void *stream) {
    // the reference's unit: one 64-sample-wide luma column of a superblock
    // (its staging buffer is 68 luma samples wide)
    if (!ctx || !dst || !src || !left || !top || !bottom || w <= 0 || h <= 0 || ss_hor < 0 || ss_hor > 1 ||
        ss_ver < 0 || ss_ver > 1 || (w << ss_hor) > 64 || (h << ss_ver) > 64 || ext_filter < 0 || ext_filter > 6 ||
        max_band_log2 < 0 || max_band_log2 > bitdepth)
        return bad("ccso_prep: invalid arguments (a luma block of at most 64 x 64)");
    if (bad_depth(bitdepth)) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    static const int8_t pos[7][2] = {{-1, 0}, {0, -1}, {-1, -1}, {-1, 1}, {-1, -2}, {1, -2}, {0, 2}};  // ccso.rs:5-13
    DspCcsoPrepArgs a;
    a.dst = dst;
    a.src = (const uint8_t *)src;
    a.left = (const uint8_t *)left;
    a.top = (const uint8_t *)top;
    a.bottom = (const uint8_t *)bottom;
    a.dst_stride = dst_stride;
    a.src_stride = src_stride;
    a.shift = bitdepth - max_band_log2;
    a.dy = pos[ext_filter][0];
    a.dx = pos[ext_filter][1];
    a.quant_step = quant_step;
    a.edge_clf = edge_clf;
    a.bo_only = bo_only != 0;
    a.edges = edges;
    a.w = w;
    a.h = h;
    a.ss_hor = ss_hor;
    a.ss_ver = ss_ver;
    R2G_DSP_LAUNCH(dsp_ccso_prep, 1, 256, pick_stream(ctx, stream), a);
    return RAV2D_OK;
}

int rav2d_gpu_ccso_add(Rav2dGpuCtx *ctx, void *dst, ptrdiff_t dst_stride, const uint8_t *idx, ptrdiff_t idx_stride,
                       const uint8_t offset_idxs[64], const int8_t offset_lut[8], int w, int h,
                       const uint16_t (*ll_mask)[4], int bitdepth, void *stream) {
    if (!ctx || !dst || !idx || !offset_idxs || !offset_lut || !ll_mask || w <= 0 || h <= 0 || w > 256 || h > 64)
        return bad("ccso_add: invalid arguments");
    if (bad_depth(bitdepth)) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    DspCcsoAddArgs a;
    memset(&a, 0, sizeof(a));
    a.dst = (uint8_t *)dst;
    a.idx = idx;
    a.dst_stride = dst_stride;
    a.idx_stride = idx_stride;
    memcpy(a.offs, offset_idxs, 64);
    memcpy(a.lut, offset_lut, 8);
    for (int r = 0; r < (h + 3) / 4; r++) a.ll[r] = ll_mask[r][0];
    a.w = w;
    a.h = h;
    a.bdmax = (1 << bitdepth) - 1;
    R2G_DSP_LAUNCH(dsp_ccso_add, 1, 256, pick_stream(ctx, stream), a);
    return RAV2D_OK;
}

int rav2d_gpu_lr_wiener(Rav2dGpuCtx *ctx, int kind, void *dst, ptrdiff_t dst_stride, const void *left,
                        const void *top, const void *bottom, int w, int h, const Rav2dWienerParams *params,
                        uint8_t edges, const uint16_t (*ll_mask)[4], int bitdepth, void *stream) {
    if (!ctx || !dst || !left || !top || !bottom || !params || !ll_mask || kind < 0 || kind > 3 || w <= 0 || h <= 0 ||
        w > 64 || h > 64 || (w & 3) || (h & 3))
        return bad("lr_wiener: invalid arguments (kind 0..3, w and h multiples of 4 up to 64)");
    if (bad_depth(bitdepth)) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    DspLrArgs *a = new DspLrArgs();
    memset(a, 0, sizeof(*a));
    a->s = stripe_of(dst, dst_stride, left, top, bottom, w, h, edges);
    a->dst = (uint8_t *)dst;
    a->kind = kind;
    a->bdmax = (1 << bitdepth) - 1;
    a->bdm8 = bitdepth - 8;
    for (int r = 0; r < h / 4 + 1; r++) a->ll[r] = ll_mask[r][0];
    if (kind <= 1) {
        const Rav2dWienerParamsSingle &p = params->single;
        if (!p.filter) {
            delete a;
            return bad("lr_wiener: no filter");
        }
        memcpy(a->filt, p.filter, kind == 0 ? 16 : 18);
        if (kind == 1) {
            if (!p.luma || !p.luma_top || !p.luma_bottom || p.ss_hor != 1 || p.ss_ver != 1 || p.ds_flt < 0 ||
                p.ds_flt > 2) {
                delete a;
                return bad("lr_wiener: chroma needs 4:2:0 luma rows and ds_flt 0..2");
            }
            a->l.luma = (const uint8_t *)p.luma;
            a->l.top = (const uint8_t *)p.luma_top;
            a->l.bottom = (const uint8_t *)p.luma_bottom;
            a->l.stride = p.stride;
            a->l.ds = p.ds_flt;
        }
    } else {
        const Rav2dWienerParamsMulti &p = params->multi;
        if (!p.filters || !p.subclass_lut || !p.noskip_mask) {
            delete a;
            return bad("lr_wiener: multi-class parameters missing");
        }
        if (kind == 2)
            memcpy(a->f.user, p.filters, sizeof(a->f.user));
        else
            memcpy(a->f.pre, p.filters, sizeof(a->f.pre));
        memcpy(a->sub, p.subclass_lut, 256);
        for (int r = 0; r < h / 4 + 1 && r < 17; r++) a->noskip[r] = p.noskip_mask[r];
        a->base_q = p.base_q;
        LrTables t;
        int rc = lr_tables(ctx->device, &t);
        if (rc) {
            delete a;
            return rc;
        }
        a->lut_to_class = t.lut_to_class;
    }
    DspLrArgs args = *a;
    delete a;
    R2G_DSP_LAUNCH(dsp_lr_wiener, 1, 256, pick_stream(ctx, stream), args);
    return RAV2D_OK;
}

int rav2d_gpu_gdf_prep(Rav2dGpuCtx *ctx, int8_t *dst, ptrdiff_t dst_stride, const void *p, ptrdiff_t stride,
                       const void *left, const void *top, const void *bottom, int w, int h, int ref_dst_idx,
                       int qp_idx, uint8_t edges, int bitdepth, void *stream) {
    if (!ctx || !dst || !p || !left || !top || !bottom || w <= 0 || h <= 0 || w > 64 || h > 64 || (w & 1) ||
        (h & 1) || ref_dst_idx < 0 || ref_dst_idx > 5 || qp_idx < 0 || qp_idx > 5)
        return bad("gdf_prep: invalid arguments");
    if (bad_depth(bitdepth)) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    LrTables t;
    int rc = lr_tables(ctx->device, &t);
    if (rc) return rc;
    DspGdfArgs a;
    a.s = stripe_of(p, stride, left, top, bottom, w, h, edges);
    a.dst = dst;
    a.dst_stride = dst_stride;
    a.up = bitdepth == 8 ? 2 : 0;
    a.down = bitdepth == 12 ? 2 : 0;
    a.alpha_base = ref_dst_idx * 528 + qp_idx * 88;
    a.weight_base = ref_dst_idx * 1584 + qp_idx * 264;
    a.bias_idx = ref_dst_idx * 6 + qp_idx;
    a.intra = ref_dst_idx == 0;
    a.err_base = a.intra ? qp_idx * 4096 : (ref_dst_idx - 1) * 6000 + qp_idx * 1000;
    a.scale = a.intra ? 8 : 5;
    a.alpha = t.gdf_alpha;
    a.weight = t.gdf_weight;
    a.bias = t.gdf_bias;
    a.err_intra = t.gdf_intra;
    a.err_inter = t.gdf_inter;
    R2G_DSP_LAUNCH(dsp_gdf_prep, 1, 256, pick_stream(ctx, stream), a);
    return RAV2D_OK;
}

int rav2d_gpu_gdf_add(Rav2dGpuCtx *ctx, void *p, ptrdiff_t stride, const int8_t *err, ptrdiff_t err_stride, int w,
                      int h, int scale, const uint16_t (*ll_mask)[4], int bitdepth, void *stream) {
    if (!ctx || !p || !err || !ll_mask || w <= 0 || h <= 0 || w > 64 || h > 64) return bad("gdf_add: invalid arguments");
    if (bad_depth(bitdepth)) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    DspGdfAddArgs a;
    memset(&a, 0, sizeof(a));
    a.p = (uint8_t *)p;
    a.err = err;
    a.stride = stride;
    a.err_stride = err_stride;
    for (int r = 0; r < (h + 3) / 4; r++) a.ll[r] = ll_mask[r][0];
    a.w = w;
    a.h = h;
    a.scale = scale;
    a.shift = 12 - bitdepth;
    a.bdmax = (1 << bitdepth) - 1;
    R2G_DSP_LAUNCH(dsp_gdf_add, 1, 256, pick_stream(ctx, stream), a);
    return RAV2D_OK;
}

static int grain_band(Rav2dGpuCtx *ctx, void *dst, const void *src, ptrdiff_t stride, const Rav2dFilmGrainData *data,
                      uint32_t seed, int pw, const uint8_t *scaling, const int16_t *grain_lut, int bh, int row_num,
                      const void *luma_row, ptrdiff_t luma_stride, int chroma, int uv, int is_id, int sx, int sy,
                      int bitdepth, void *stream) {
    if (!ctx || !dst || !src || !data || !scaling || !grain_lut || pw <= 0 || pw > RAV2D_MAX_FRAME_DIM || bh <= 0 ||
        row_num < 0 || data->block_size < 0 || data->block_size > 1 || data->scaling_shift < 8 ||
        data->scaling_shift > 11 || bh > (16 << data->block_size) ||
        (chroma && (!luma_row || sx < sy || sx < 0 || sx > 1 || sy < 0 || uv < 0 || uv > 1)))
        return bad("film grain band: invalid arguments");
    if (bad_depth(bitdepth)) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    DspGrainArgs a;
    memset(&a, 0, sizeof(a));
    a.dst = (uint8_t *)dst;
    a.src = (const uint8_t *)src;
    a.luma = (const uint8_t *)luma_row;
    a.stride = stride;
    a.luma_stride = luma_stride;
    a.scaling = scaling;
    a.lut = grain_lut;
    a.fg = *data;
    a.seed = seed;
    a.pw = pw;
    a.bh = bh;
    a.row_num = row_num;
    a.bitdepth = bitdepth;
    a.chroma = chroma;
    a.uv = uv;
    a.is_id = is_id != 0;
    a.sx = chroma ? sx : 0;
    a.sy = chroma ? sy : 0;
    R2G_DSP_LAUNCH(dsp_grain_band, (pw + 63) / 64, 256, pick_stream(ctx, stream), a);
    return RAV2D_OK;
}

static int grain_params_ok(const Rav2dFilmGrainData *d) {
    return d && d->ar_coeff_lag >= 0 && d->ar_coeff_lag <= 3 && d->ar_coeff_shift >= 6 && d->ar_coeff_shift <= 9 &&
           d->grain_scale_shift >= 0 && d->grain_scale_shift <= 3;
}

int rav2d_gpu_generate_grain_y(Rav2dGpuCtx *ctx, int16_t *buf, const Rav2dFilmGrainData *data, uint32_t seed,
                               int bitdepth, void *stream) {
    if (!ctx || !buf || !grain_params_ok(data)) return bad("generate_grain_y: invalid arguments");
    if (bad_depth(bitdepth)) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    return launch_grain_generate(*data, seed, bitdepth, 0, 0, 0, buf, nullptr, pick_stream(ctx, stream));
}

int rav2d_gpu_generate_grain_uv(Rav2dGpuCtx *ctx, int16_t *buf, const int16_t *buf_y,
                                const Rav2dFilmGrainData *data, uint32_t seed, int uv, int subx, int suby,
                                int bitdepth, void *stream) {
    // (subx, suby): (1, 1) 4:2:0, (1, 0) 4:2:2, (0, 0) 4:4:4
    if (!ctx || !buf || !buf_y || !grain_params_ok(data) || uv < 0 || uv > 1 || subx < suby || subx < 0 || subx > 1 ||
        suby < 0)
        return bad("generate_grain_uv: invalid arguments");
    if (bad_depth(bitdepth)) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    return launch_grain_generate(*data, seed, bitdepth, 1 + uv, subx, suby, buf, buf_y, pick_stream(ctx, stream));
}

int rav2d_gpu_fgy_32x32xn(Rav2dGpuCtx *ctx, void *dst, const void *src, ptrdiff_t stride,
                          const Rav2dFilmGrainData *data, uint32_t seed, int pw, const uint8_t *scaling,
                          const int16_t *grain_lut, int bh, int row_num, int bitdepth, void *stream) {
    return grain_band(ctx, dst, src, stride, data, seed, pw, scaling, grain_lut, bh, row_num, nullptr, 0, 0, 0, 0, 0,
                      0, bitdepth, stream);
}

int rav2d_gpu_fguv_32x32xn(Rav2dGpuCtx *ctx, void *dst, const void *src, ptrdiff_t stride,
                           const Rav2dFilmGrainData *data, uint32_t seed, int pw, const uint8_t *scaling,
                           const int16_t *grain_lut, int bh, int row_num, const void *luma_row,
                           ptrdiff_t luma_stride, int uv, int is_id, int sx, int sy, int bitdepth, void *stream) {
    return grain_band(ctx, dst, src, stride, data, seed, pw, scaling, grain_lut, bh, row_num, luma_row, luma_stride, 1,
                      uv, is_id, sx, sy, bitdepth, stream);
}

}  // extern "C"
