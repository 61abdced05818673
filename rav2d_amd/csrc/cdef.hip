// CDEF with CCSO on gfx950: one workgroup per 64x64 luma superblock (and its
// two 32x32 chroma blocks), F2 -> F3.
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

__constant__ int8_t kCdefDyDx[12][2][2];  // CDEF_DIRECTIONS as (dy, dx)
__constant__ int8_t kCcsoOffset[4][8];

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

// Direction search of one 8x8 luma block staged at `img` (row stride `ls`).
// The eight line families are accumulated in two rounds of four (re-reading
// the 64 samples from LDS), which halves the live partial sums.
__device__ int find_dir(const int16_t *img, int ls, uint32_t *var, int bdm8) {
    // partial sums are at most 8 x 128 in magnitude: the squares are 24-bit products
    auto sq = [](int v) -> uint32_t { return (uint32_t)((int)(int16_t)v * (int)(int16_t)v); };
    const uint32_t DIV[7] = {840, 420, 280, 210, 168, 140, 120};
    uint32_t cost[8];
    {
        int hv0[8] = {0}, hv1[8] = {0}, d0[15] = {0}, d1[15] = {0};
#pragma unroll
        for (int y = 0; y < 8; y++) {
#pragma unroll
            for (int x = 0; x < 8; x++) {
                int px = ((int)img[y * ls + x] >> bdm8) - 128;
                d0[y + x] += px;
                hv0[y] += px;
                d1[7 + y - x] += px;
                hv1[x] += px;
            }
            __builtin_amdgcn_sched_barrier(0);  // one row of loads in flight at a time
        }
        cost[2] = 0;
        cost[6] = 0;
#pragma unroll
        for (int n = 0; n < 8; n++) {
            cost[2] += sq(hv0[n]);
            cost[6] += sq(hv1[n]);
        }
        cost[2] *= 105;
        cost[6] *= 105;
        cost[0] = 0;
        cost[4] = 0;
#pragma unroll
        for (int n = 0; n < 7; n++) {
            cost[0] += (sq(d0[n]) + sq(d0[14 - n])) * DIV[n];
            cost[4] += (sq(d1[n]) + sq(d1[14 - n])) * DIV[n];
        }
        cost[0] += sq(d0[7]) * 105;
        cost[4] += sq(d1[7]) * 105;
    }
    {
        int a0[11] = {0}, a1[11] = {0}, a2[11] = {0}, a3[11] = {0};
#pragma unroll
        for (int y = 0; y < 8; y++) {
#pragma unroll
            for (int x = 0; x < 8; x++) {
                int px = ((int)img[y * ls + x] >> bdm8) - 128;
                a0[y + (x >> 1)] += px;
                a1[3 + y - (x >> 1)] += px;
                a2[3 - (y >> 1) + x] += px;
                a3[(y >> 1) + x] += px;
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        int *alts[4] = {a0, a1, a2, a3};
#pragma unroll
        for (int n = 0; n < 4; n++) {
            int *al = alts[n];
            uint32_t c = 0;
#pragma unroll
            for (int m = 0; m < 5; m++) c += sq(al[3 + m]);
            c *= 105;
#pragma unroll
            for (int m = 0; m < 3; m++) c += (sq(al[m]) + sq(al[10 - m])) * DIV[2 * m + 1];
            cost[n * 2 + 1] = c;
        }
    }
    int best_dir = 0;
    uint32_t best = cost[0];
#pragma unroll
    for (int n = 1; n < 8; n++)
        if (cost[n] > best) {
            best = cost[n];
            best_dir = n;
        }
    *var = (best - cost[best_dir ^ 4]) >> 10;
    return best_dir;
}

// Two horizontally adjacent output samples of cdef_filter_block in packed
// 16-bit lanes.  `t` points at the left sample in an LDS tile whose
// outside-frame entries hold INT16_MIN; `op`, `os0`, `os1` are the tap offsets
// (in that tile's stride) of the primary direction and of the two secondary
// directions, for taps k = 0, 1.  16 bits hold every intermediate: tap
// differences are formed with a saturating subtract, so a difference with an
// INT16_MIN neighbour stays at least 32767 in magnitude and constrains to 0
// exactly as in 32-bit arithmetic (thresholds are at most 15 << 2 and shifts
// at most 8); |sum| <= 12 * 4 * 60 < 2^15.
typedef short s16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 splat2(int v) { return (s16x2){(short)v, (short)v}; }

__device__ __forceinline__ s16x2 constrain2(s16x2 d, s16x2 thr, s16x2 sh) {
    const s16x2 z = {0, 0};
    const s16x2 ad = __builtin_elementwise_max(d, __builtin_elementwise_sub_sat(z, d));
    const s16x2 c = __builtin_elementwise_max(z, thr - (ad >> sh));
    return __builtin_elementwise_min(__builtin_elementwise_max(d, z - c), c);
}

__device__ __forceinline__ s16x2 cdef_px2(const int16_t *t, const int op[2], const int os0[2], const int os1[2],
                                          int pri, int pri_tap, int pri_shift, int sec, int sec_shift) {
    // two 16-bit LDS reads per pair: odd tap offsets would make a merged 32-bit
    // read unaligned (volatile keeps the compiler from merging them)
    const volatile int16_t *vt = t;
    auto ld = [&](int o) -> s16x2 { return (s16x2){vt[o], vt[o + 1]}; };
    const s16x2 px = ld(0);
    const s16x2 ps = splat2(pri), pss = splat2(pri_shift), ss = splat2(sec), sss = splat2(sec_shift);
    auto dif = [&](s16x2 v) { return __builtin_elementwise_sub_sat(v, px); };
    s16x2 sum = {0, 0};
    if (pri) {
        s16x2 tk = splat2(pri_tap);
        if (sec) {
            s16x2 mx = px, mn = px;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const s16x2 p0 = ld(op[k]), p1 = ld(-op[k]);
                sum += tk * constrain2(dif(p0), ps, pss);
                sum += tk * constrain2(dif(p1), ps, pss);
                tk = (tk & (short)3) | (short)2;
                mn = __builtin_elementwise_min(mn, __builtin_elementwise_min(p0, p1));
                mx = __builtin_elementwise_max(mx, __builtin_elementwise_max(p0, p1));
                const s16x2 s0 = ld(os0[k]), s1 = ld(-os0[k]), s2 = ld(os1[k]), s3 = ld(-os1[k]);
                const s16x2 st = splat2(2 - k);
                sum += st * constrain2(dif(s0), ss, sss);
                sum += st * constrain2(dif(s1), ss, sss);
                sum += st * constrain2(dif(s2), ss, sss);
                sum += st * constrain2(dif(s3), ss, sss);
                mn = __builtin_elementwise_min(mn, __builtin_elementwise_min(__builtin_elementwise_min(s0, s1),
                                                                              __builtin_elementwise_min(s2, s3)));
                mx = __builtin_elementwise_max(mx, __builtin_elementwise_max(__builtin_elementwise_max(s0, s1),
                                                                              __builtin_elementwise_max(s2, s3)));
            }
            const s16x2 v = px + ((sum + (sum >> (short)15) + (short)8) >> (short)4);
            return __builtin_elementwise_min(__builtin_elementwise_max(v, mn), mx);
        }
#pragma unroll
        for (int k = 0; k < 2; k++) {
            sum += tk * constrain2(dif(ld(op[k])), ps, pss);
            sum += tk * constrain2(dif(ld(-op[k])), ps, pss);
            tk = (tk & (short)3) | (short)2;
        }
        return px + ((sum + (sum >> (short)15) + (short)8) >> (short)4);
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const s16x2 st = splat2(2 - k);
        sum += st * constrain2(dif(ld(os0[k])), ss, sss);
        sum += st * constrain2(dif(ld(-os0[k])), ss, sss);
        sum += st * constrain2(dif(ld(os1[k])), ss, sss);
        sum += st * constrain2(dif(ld(-os1[k])), ss, sss);
    }
    return px + ((sum + (sum >> (short)15) + (short)8) >> (short)4);
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
    int16_t off[2][12][2]; // CDEF_DIRECTIONS as offsets in the luma / chroma tile
};

// CCSO offset of one sample: `lt` = staged luma at the collocated position
// (lx, lyy in frame coordinates), neighbours clamped to the frame.
__device__ __forceinline__ int ccso_offset(const CcsoDev &c, const int16_t *ly, int lx, int lyy, int x0, int y0, int fw,
                                           int fh) {
    const int cval = ly[(lyy - y0 + 2) * LY + lx - x0 + CX];
    int idx = (int)((uint32_t)cval >> c.shift);
    if (!c.bo_only) {
        int nx0 = iclip(lx + c.dx, 0, fw - 1) - x0 + CX, ny0 = iclip(lyy + c.dy, 0, fh - 1) - y0 + 2;
        int nx1 = iclip(lx - c.dx, 0, fw - 1) - x0 + CX, ny1 = iclip(lyy - c.dy, 0, fh - 1) - y0 + 2;
        uint32_t c0 = ccso_score(ly[ny0 * LY + nx0] - cval, c.quant_step, c.edge_clf);
        uint32_t c1 = ccso_score(ly[ny1 * LY + nx1] - cval, c.quant_step, c.edge_clf);
        idx |= (int)((c0 << 5) | (c1 << 3));
        idx &= 0xff;
    }
    const int packed = (idx >> 1) < 64 ? c.filter_off[idx >> 1] : 0;
    return kCcsoOffset[c.scale_idx][7 & (packed >> (4 * (idx & 1)))];
}

template <typename P>
__device__ __forceinline__ void cdef_body(const CdefArgs &a, int sbx, int sby, CdefSmem &sm) {
    int16_t *ly = sm.ly;
    int16_t(*lc)[LCR * LC] = sm.lc;
    uint8_t *b_state = sm.b_state, *b_dir = sm.b_dir;
    int16_t *b_ypri = sm.b_ypri;
    const int t = threadIdx.x;
    const int x0 = sbx * 64, y0 = sby * 64;
    const int fw = a.g.bw * 4, fh = a.g.bh * 4;
    const int cw = fw >> 1, ch = fh >> 1;
    const int bdm8 = a.g.bitdepth - 8;
    // ---- staging of F2 with INT16_MIN outside the frame; every load of a lane
    // is issued before its first LDS store
    {
        // 4-sample groups (tile origins and plane widths are multiples of 4, so a
        // group is wholly inside or outside the frame), loaded from clamped,
        // always valid addresses, then selected
        constexpr int QY = LY / 4, QC = LC / 4;
        constexpr int NY = LYR * QY, NC = 2 * LCR * QC;
        constexpr int PY = (NY + kThreads - 1) / kThreads, PC = (NC + kThreads - 1) / kThreads;
        const uint2 outside = make_uint2(0x80008000u, 0x80008000u);  // INT16_MIN x 4
        uint2 vy[PY], vc[PC];
#pragma unroll
        for (int j = 0; j < PY; j++) {
            const int i = t + j * kThreads;
            const int r = i / QY, gx = x0 - CX + (i - r * QY) * 4, gy = y0 - 2 + r;
            const uint2 v = load4(row_ptr<const P>(a.src[0], iclip(gy, 0, fh - 1)) + iclip(gx, 0, fw - 4));
            vy[j] = (gx >= 0 && gx < fw && gy >= 0 && gy < fh) ? v : outside;
        }
#pragma unroll
        for (int j = 0; j < PC; j++) {
            const int i = t + j * kThreads;
            const int pl = i >= LCR * QC;
            const int k = i - pl * LCR * QC;
            const int r = k / QC, gx = (x0 >> 1) - CX + (k - r * QC) * 4, gy = (y0 >> 1) - 2 + r;
            const uint2 v =
                load4(row_ptr<const P>(a.src[1 + (pl & 1)], iclip(gy, 0, ch - 1)) + iclip(gx, 0, cw - 4));
            vc[j] = (gx >= 0 && gx < cw && gy >= 0 && gy < ch) ? v : outside;
        }
#pragma unroll
        for (int j = 0; j < PY; j++)
            if (t + j * kThreads < NY) ((uint2 *)ly)[t + j * kThreads] = vy[j];
#pragma unroll
        for (int j = 0; j < PC; j++)
            if (t + j * kThreads < NC) ((uint2 *)&lc[0][0])[t + j * kThreads] = vc[j];
        if (t < 48) {
            const int tab = t / 24, d = (t >> 1) % 12, k = t & 1;
            sm.off[tab][d][k] = (int16_t)(kCdefDyDx[d][k][0] * (tab ? LC : LY) + kCdefDyDx[d][k][1]);
        }
    }
    const Rav2dAv2Filter &m = a.mask[(sby >> 2) * a.g.sb256w + (sbx >> 2)];
    const int sb64_idx = ((sby & 3) << 2) + (sbx & 3);
    const int cdef_idx = m.cdef_idx[sb64_idx];
    const bool active = a.cdef_on && cdef_idx >= 0 &&
                        (a.y_strength[cdef_idx] != 0 || a.uv_strength[cdef_idx] != 0);
    const int y_lvl = active ? a.y_strength[cdef_idx] : 0;
    const int uv_lvl = active ? a.uv_strength[cdef_idx] : 0;
    const int y_pri = (y_lvl >> 2) << bdm8;
    int y_sec = y_lvl & 3;
    y_sec += y_sec == 3;
    y_sec <<= bdm8;
    const int uv_pri = (uv_lvl >> 2) << bdm8;
    int uv_sec = uv_lvl & 3;
    uv_sec += uv_sec == 3;
    uv_sec <<= bdm8;
    const int damping = a.damping + bdm8;
    __syncthreads();
    if (t < 64) {
        int bxi = t & 7, byi = t >> 3;
        int bx = (x0 >> 2) + bxi * 2, by = (y0 >> 2) + byi * 2;  // 4px units
        uint8_t st = 0, dir = 0;
        int16_t ypri = 0;
        if (active && bx < a.g.bw && by < a.g.bh) {
            uint16_t noskip = a.on_skiptx ? 0xffff : m.noskip_mask[(by & 0x3e) >> 1][sbx & 3];
            if (noskip & (3u << (bx & 14))) {
                uint32_t var = 0;
                int d = (y_pri || uv_pri) ? find_dir(&ly[(byi * 8 + 2) * LY + bxi * 8 + CX], LY, &var, bdm8) : 0;
                if (y_pri) {
                    int adj = adjust_strength(y_pri, var);
                    if (adj || y_sec) st |= 1;
                    ypri = (int16_t)adj;
                } else if (y_sec) {
                    st |= 1;
                }
                if (uv_lvl) st |= 2;
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
        const bool cc0 = a.ccso_on && m.ccso[0];
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
                if (cc0) {
                    v0 = iclip(v0 + ccso_offset(a.ccso[0], ly, gx, gy, x0, y0, fw, fh), 0, a.g.bdmax);
                    v1 = iclip(v1 + ccso_offset(a.ccso[0], ly, gx + 1, gy, x0, y0, fw, fh), 0, a.g.bdmax);
                }
                store2(row_ptr<P>(a.dst[0], gy) + gx, v0, v1);
            }
        }
    }
    // ---- chroma: lane t filters the column pair 2 (t & 15), +1 of the 4 rows of
    // block row (t >> 4) & 7 of plane t >> 7 (one 4x4 block per lane)
    {
        const int pl = t >> 7, x = 2 * (t & 15), by = 4 * ((t >> 4) & 7);
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
        const bool cc = a.ccso_on && m.ccso[1 + pl];
        if (gx < cw) {
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
                if (cc) {
                    v0 = iclip(v0 + ccso_offset(a.ccso[1 + pl], ly, gx << 1, gy << 1, x0, y0, fw, fh), 0,
                               a.g.bdmax);
                    v1 = iclip(v1 + ccso_offset(a.ccso[1 + pl], ly, (gx + 1) << 1, gy << 1, x0, y0, fw, fh), 0,
                               a.g.bdmax);
                }
                store2(row_ptr<P>(a.dst[1 + pl], gy) + gx, v0, v1);
            }
        }
    }
}

template <typename P>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6, 8))) void cdef_kernel(CdefArgs a) {
    __shared__ CdefSmem sm;
    cdef_body<P>(a, blockIdx.x, blockIdx.y, sm);
}

// Batched launch: one item per 64x64 superblock of every frame of one bit depth.
template <typename P, bool kLoop>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6, 8))) void cdef_batch(const CdefArgs *__restrict__ args,
                                                                                        BatchSched s) {
    __shared__ CdefSmem sm;
    for_each_item<kLoop>(s, [&](int f, int t) {
        const CdefArgs &a = args[f];
        const int sbw = (a.g.bw * 4 + 63) >> 6;
        const int sby = t / sbw, sbx = t - sby * sbw;
        cdef_body<P>(a, sbx, sby, sm);
        __syncthreads();
    });
}

int cdef_items(const CdefArgs &a) { return ((a.g.bw * 4 + 63) >> 6) * ((a.g.bh * 4 + 63) >> 6); }

// Per device: the direction offsets as (dy, dx) and the CCSO offset table.
void cdef_init_consts() {
    static bool done[64] = {};
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64 || done[dev]) return;
    {
        int8_t dydx[12][2][2];
        for (int d = 0; d < 12; d++)
            for (int k = 0; k < 2; k++) {
                int o = R2T_CDEF_DIRECTIONS[d][k];  // dy * 12 + dx with |dx| <= 2
                int dx = ((o + 2) % 12 + 12) % 12 - 2;
                dydx[d][k][0] = (int8_t)((o - dx) / 12);
                dydx[d][k][1] = (int8_t)dx;
            }
        (void)hipMemcpyToSymbol(HIP_SYMBOL(kCdefDyDx), dydx, sizeof(dydx));
        (void)hipMemcpyToSymbol(HIP_SYMBOL(kCcsoOffset), R2T_CCSO_OFFSET, sizeof(R2T_CCSO_OFFSET));
        done[dev] = true;
    }
}

void launch_cdef_batch(const CdefArgs *d_args, const BatchSched &s, int grid, int bitdepth, hipStream_t st) {
    if (s.total <= 0) return;
    const bool loop = grid < s.total;
    if (bitdepth > 8) {
        if (loop)
            hipLaunchKernelGGL((cdef_batch<uint16_t, true>), dim3(grid), dim3(kThreads), 0, st, d_args, s);
        else
            hipLaunchKernelGGL((cdef_batch<uint16_t, false>), dim3(grid), dim3(kThreads), 0, st, d_args, s);
    } else {
        if (loop)
            hipLaunchKernelGGL((cdef_batch<uint8_t, true>), dim3(grid), dim3(kThreads), 0, st, d_args, s);
        else
            hipLaunchKernelGGL((cdef_batch<uint8_t, false>), dim3(grid), dim3(kThreads), 0, st, d_args, s);
    }
}

void launch_cdef(const CdefArgs &a, int bitdepth, hipStream_t s) {
    cdef_init_consts();
    dim3 grid((a.g.bw * 4 + 63) / 64, (a.g.bh * 4 + 63) / 64);
    if (bitdepth > 8)
        hipLaunchKernelGGL(cdef_kernel<uint16_t>, grid, dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL(cdef_kernel<uint8_t>, grid, dim3(kThreads), 0, s, a);
}

}  // namespace r2g
