// Film-grain synthesis on gfx950: output planes -> display planes.
//
// Arithmetic follows crates/rav2d/src/filmgrain.rs: the 16-bit LFSR
// get_random_number (:11-16), round2 (:18-20), generate_scaling_8bpc/hbd
// (:22-113), generate_grain_y/uv (:115-229), sample_lut (:231-246) and the
// per-block overlap blending of fgy_32x32xn / fguv_32x32xn (:248-592), driven
// per row band as apply_grain_row / apply_grain_mt (:688-941).
//
// Two kernels.  grain_setup (one workgroup) builds everything that depends
// only on the frame's grain parameters: the three grain LUTs (the Gaussian
// draws are one LFSR walk per plane; the causal AR filter runs as a wavefront
// over anti-diagonals x + 4y, since a sample needs its three left neighbours
// and the rows above up to x+3), the scaling LUTs, and the per-(row band,
// block) LUT offsets (each band's LFSR walk, by one lane per band; a band's
// "previous row" offsets are the previous band's own).  grain_apply then
// computes every output sample independently.
#include <string.h>

#include <mutex>

#include "internal.h"
#include "tables.h"

namespace r2g {

constexpr int GW = 82, GH = 73, SGW = 44, SGH = 38;

__device__ __forceinline__ uint32_t lfsr(uint32_t bits, uint32_t &s) {
    uint32_t r = s;
    uint32_t bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1;
    s = (r >> 1) | (bit << 15);
    return (s >> (16 - bits)) & ((1u << bits) - 1);
}

__device__ __forceinline__ int round2(int x, int shift) { return (x + ((1 << shift) >> 1)) >> shift; }

// LFSR jump-ahead: the 16-bit LFSR is linear over GF(2), so its state n
// draws on is M^n applied to the state.  Jump tables hold such matrices as 16
// row masks (output bit i = parity(row_i & state)); the host builds them once
// per device.  kJump[k] = M^(k kChunk) starts lane k's share of a plane's
// Gaussian draws; kJumpSeg[j] = M^(8 kBandSeg j) starts segment j of a row
// band's block offsets (each block takes 8 draws).
constexpr int kChunk = 47, kBandSeg = 16, kSegs = 128;
__constant__ uint16_t kJump[128][16];
__constant__ uint16_t kJumpSeg[kSegs][16];

__device__ __forceinline__ uint32_t jump(const uint16_t (&m)[16], uint32_t s) {
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) r |= (uint32_t)(__popc(m[i] & s) & 1) << i;
    return r;
}

// four waves: set-up phases run the luma, U and V AR filters and the scaling
// LUTs / block offsets side by side
constexpr int kSetupThreads = 256;

struct GrainSmem {
    int16_t lut[3][GH][GW];
    int16_t g_lds[2048];
    uint8_t scl[3][4096];  // 1 << bitdepth entries per plane
    int8_t win[3][4][7];
    int luma_done;  // luma AR steps completed, published per 16-step block
};

// Causal AR filter of one grain plane by one wave, as a wavefront whose
// recent samples stay in registers.  Lane k owns row k (for the luma plane
// also row 64 + k, k < 9, from step 128: rows k end by step 4k + 81 < 128 and
// rows 64 + k start at step 256 + 4k).  Sample (y, x) is final at step x + 4y:
// at step s every lane emits the value of (y, s - 4y), filtered inside the AR
// area and the Gaussian draw outside it, and keeps four 16-entry rings indexed
// by step: E its own emissions, A lane k-1's (row y - 1, one step late), B
// row y - 2's (two steps late, relayed by lane k-1), C row y - 3's (three
// steps late).  Columns x-3..x+3 of row y - 1 were emitted at steps s-7..s-1,
// of row y - 2 at s-11..s-5, of row y - 3 at s-15..s-9, and the own row's
// x-3..x-1 at s-3..s-1, so each step needs only its three lane rotations (DPP
// wave_ror:1, lane k reads lane k-1) and no LDS round trip on the dependency
// chain.  Taps: win[r][c] for dy = r - 3, dx = c - 3 (zero outside the lag).
//
// The luma wave publishes its completed step count in *done after every
// block of 16 steps (after its LDS writes).  A chroma sample (y, x) at step
// c = x + 4y reads luma rows 2y-3, 2y-2 and columns 2x-3, 2x-2, final by luma
// step 2c - 10, so a chroma wave runs alongside the luma one and waits, before
// its block b, for 32b + 21 luma steps (when it uses the luma term).
typedef short s16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int lane_up(int v) { return __builtin_amdgcn_mov_dpp(v, 0x13C, 0xF, 0xF, true); }

constexpr int kLumaSteps = (GW - 1) + 4 * (GH - 1) + 1;

// kChroma: a chroma plane of subsampling (SX, SY) (template (SY ? 38 : 73) x
// (SX ? 44 : 82)), whose AR filter adds the average of the co-located luma
// template samples; 4:2:0 chroma runs alongside the luma wave, 4:2:2 / 4:4:4
// chroma waits for the whole luma template.
template <bool kChroma, int SX = 1, int SY = 1>
__device__ __forceinline__ void ar_wave(int16_t (*L)[GW], const int8_t (*win)[7], int ar_shift, int gmin, int gmax,
                                        const int16_t (*luma)[GW], int wl, int *done, uint32_t *err) {
    constexpr int W = kChroma ? (SX ? SGW : GW) : GW, H = kChroma ? (SY ? SGH : GH) : GH;
    constexpr int kSteps = (W - 1) + 4 * (H - 1) + 1;
    constexpr int kLumaPublished = 16 * ((kLumaSteps + 15) / 16);
    static_assert(!kChroma || !(SX && SY) || 32 * ((kSteps + 15) / 16 - 1) + 21 <= kLumaPublished,
                  "every chroma block's luma steps get published");
    static_assert(4 * 8 + W - 1 < 128 && (H <= 64 || H - 64 <= 9), "rows 64.. start after the switch");
    const int lane = threadIdx.x & 63;
    int cf[4][7];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 7; c++) cf[r][c] = __builtin_amdgcn_readfirstlane((int)win[r][c]);
    int E[16], A[16], B[16], C[16];
#pragma unroll
    for (int i = 0; i < 16; i++) E[i] = A[i] = B[i] = C[i] = 0;
    const int16_t *Lf = &L[0][0];
    const int16_t *lf = kChroma ? &luma[0][0] : nullptr;
    const int rnd = (1 << ar_shift) >> 1;
    int y = lane;
    for (int b = 0; b < (kSteps + 15) / 16; b++) {
        if (H > 64 && b == 8 && lane < H - 64) y = lane + 64;
        if (kChroma && wl != 0) {
            // bounded wait (the luma wave never waits, so the count always
            // arrives); should it not, the error word tells the host, which then
            // fails the frame's next synchronizing call instead of returning
            // grain built from unfinished luma rows
            const int need = (SX && SY) ? 32 * b + 21 : kLumaPublished;
            bool arrived = false;
            for (int spin = 0; spin < (1 << 20); spin++) {
                if (__hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= need) {
                    arrived = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (!arrived && lane == 0 && err) atomicOr(err, kGrainErrLumaWait);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        const bool rows_ok = y >= 3 && y < H;
        // (y, s - 4y) is element y (GW - 4) + s of the plane's rows; lanes
        // without a sample at step s read (and never write) other LDS entries
        const int base = y * (GW - 4), xm = -4 * y - 3;
        // co-located luma of chroma sample (y, x) = step s - 4y: rows ((y - 3) << SY) + 3
        // (and the next with SY), columns ((x - 3) << SX) + 3 (and the next with SX)
        const int lbase = !kChroma ? 0
                          : (SX && SY) ? dmax(y * (2 * GW - 8) - 3 * GW - 3, 0)
                          : SX         ? dmax(y * (GW - 8) - 3, 0)
                                       : y * (GW - 4);
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int s = b * 16 + u;
            A[u] = lane_up(E[(u - 1) & 15]);
            B[u] = lane_up(A[(u - 1) & 15]);
            C[u] = lane_up(B[(u - 1) & 15]);
            const int raw = Lf[base + s];
            int lt = 0;
            if constexpr (kChroma) {
                if constexpr (SX && SY) {
                    const int i = lbase + 2 * s;
                    lt = __mul24((lf[i] + lf[i + 1] + lf[i + GW] + lf[i + GW + 1] + 2) >> 2, wl);
                } else if constexpr (SX) {
                    const int i = lbase + 2 * s;
                    lt = __mul24((lf[i] + lf[i + 1] + 1) >> 1, wl);
                } else {
                    lt = __mul24((int)lf[lbase + s], wl);
                }
            }
            // four partial sums: rows y-3, y-2, y-1 and the own row
            int a0 = 0, a1 = 0, a2 = 0, a3 = lt;
#pragma unroll
            for (int c = 0; c < 7; c++) {
                a0 += __mul24(C[(u - 12 + c) & 15], cf[0][c]);
                a1 += __mul24(B[(u - 9 + c) & 15], cf[1][c]);
                a2 += __mul24(A[(u - 6 + c) & 15], cf[2][c]);
            }
#pragma unroll
            for (int c = 0; c < 3; c++) a3 += __mul24(E[(u - 3 + c) & 15], cf[3][c]);
            const int sum = (a0 + a1) + (a2 + a3);
            const bool act = rows_ok && (unsigned)(s + xm) <= (unsigned)(W - 7);
            int v = raw;
            if (act) {
                v = dmin(dmax(raw + ((sum + rnd) >> ar_shift), gmin), gmax);
                L[0][base + s] = (int16_t)v;
            }
            E[u] = v;
        }
        if constexpr (!kChroma) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (lane == 0) __hip_atomic_store(done, 16 * (b + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
}

// Scaling LUTs (generate_scaling_8bpc/hbd, filmgrain.rs:22-113) by one wave,
// every entry in closed form: the first pass writes entries at multiples of
// 1 << (bd-8) (later segments overwrite earlier ones, then the tail fill), the
// second pass interpolates between them inside each segment.
__device__ __forceinline__ void scaling_luts(const Rav2dFilmGrainData &d, int bd, uint8_t (*scl)[4096], int lane) {
    for (int pl = 0; pl < 3; pl++) {
        const int num = d.num_points[pl];
        const uint8_t(*pts)[2] = d.points[pl];
        const int sx = bd - 8, size = 1 << bd, pad = 1 << sx;
        auto first_pass = [&](int c) -> int {  // value at coarse position c
            int v = 0;
            if (c < pts[0][0]) v = pts[0][1];
            for (int k = 0; k < num - 1; k++) {
                int bx = pts[k][0], dx = pts[k + 1][0] - bx;
                if (dx > 0 && c >= bx && c < bx + dx) {
                    int delta = (pts[k + 1][1] - pts[k][1]) * ((0x10000 + (dx >> 1)) / dx);
                    v = (uint8_t)(pts[k][1] + ((0x8000 + (c - bx) * delta) >> 16));
                }
            }
            if (c >= pts[num - 1][0]) v = pts[num - 1][1];
            return v;
        };
        for (int e = lane; e < size; e += 64) {
            int v = 0;
            if (num > 0) {
                int c = e >> sx, n = e & (pad - 1);
                v = first_pass(c);
                if (n) {
                    // inside a segment (dx > 0): the second pass interpolates
                    // between the coarse entries around it, the same values
                    // whichever segment writes last; elsewhere the first
                    // pass's fills hold (points that repeat or go back, which
                    // the reference tolerates by skipping dx <= 0 segments)
                    bool in_seg = false;
                    for (int k = 0; k < num - 1; k++) {
                        int bx = pts[k][0] << sx, ex = pts[k + 1][0] << sx;
                        if (e >= bx && e < ex) in_seg = true;
                    }
                    if (in_seg) {
                        int base = v, range = first_pass(c + 1) - base;
                        v = (uint8_t)(base + (((pad >> 1) + n * range) >> sx));
                    }
                }
            }
            scl[pl][e] = (uint8_t)v;
        }
    }
}

// LUT offsets of every (row band, block): each band's LFSR walk (fgy_32x32xn,
// filmgrain.rs:276-304: the band seed, then per block and direction one 9-bit
// draw and three 16-bit ones) split into segments of kBandSeg blocks, one lane
// per (band, segment), each starting from the band seed moved on by kJumpSeg.
__device__ __forceinline__ void band_offsets(const GrainArgs &a, int lane) {
    const int nseg = (a.n_blocks + kBandSeg - 1) / kBandSeg;
    const int bs_mul = 3 - a.fg.block_size;
    for (int task = lane; task < a.n_bands * nseg; task += 64) {
        const int band = task / nseg, seg = task - band * nseg;
        uint32_t s = a.seed;
        s ^= (uint32_t)(((band * 37 + 178) & 0xFF)) << 8;
        s ^= (uint32_t)((band * 173 + 105) & 0xFF);
        s = jump(kJumpSeg[dmin(seg, kSegs - 1)], s & 0xffff);
        for (int k = kSegs - 1; k < seg; k++)  // frames wider than kSegs segments
            for (int n = 0; n < 8 * kBandSeg; n++) lfsr(16, s);
        const int b0 = seg * kBandSeg, b1 = dmin(b0 + kBandSeg, a.n_blocks);
        uint8_t *o = a.offsets + (size_t)band * a.n_blocks * 2;
        for (int b = b0; b < b1; b++)
            for (int n = 0; n < 2; n++) {
                o[b * 2 + n] = (uint8_t)(((uint32_t)bs_mul * lfsr(9, s)) >> 6);
                lfsr(16, s);
                lfsr(16, s);
                lfsr(16, s);
            }
    }
}

__device__ __forceinline__ void grain_setup_body(const GrainArgs &a, const int16_t *gauss, GrainSmem &sm) {
    auto &lut = sm.lut;
    auto &g_lds = sm.g_lds;
    auto &win = sm.win;
    const Rav2dFilmGrainData &d = a.fg;
    const int bd = a.bitdepth, bdm8 = bd - 8;
    const int shift = 4 - bdm8 + d.grain_scale_shift;
    const int ctr = 128 << bdm8, gmin = -ctr, gmax = ctr - 1;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const bool want_u = d.num_points[1] > 0 || d.chroma_scaling_from_luma;
    const bool want_v = d.num_points[2] > 0 || d.chroma_scaling_from_luma;
    const int lag = d.ar_coeff_lag;
    for (int i = t; i < 2048; i += kSetupThreads) g_lds[i] = gauss[i];
    // AR taps on a fixed lag-3 window (zero where the frame's lag is smaller):
    // window row r = dy + 3 (dy = -3..0), column c = dx + 3 (dx = -3..3); the
    // reference enumerates rows dy = -lag..0 and columns dx = -lag..lag
    // (-lag..-1 on row 0): coefficient (dy + lag) (2 lag + 1) + (dx + lag)
    if (t < 3 * 28) {
        int pl = t / 28, r = (t % 28) / 7, c = t % 7;
        int dy = r - 3, dx = c - 3;
        int v = 0;
        bool in = dy >= -lag && dx >= -lag && (dy < 0 ? dx <= lag : dx < 0);
        if (in) v = d.ar_coeffs[pl][(dy + lag) * (2 * lag + 1) + (dx + lag)];
        win[pl][r][c] = (int8_t)v;
    }
    if (t == 0) sm.luma_done = 0;
    __syncthreads();
    // Gaussian draws, kChunk per job: jobs 0..127 the luma sequence (73x82), then
    // the U and V sequences (chroma templates cw x ch: 44x38 for 4:2:0, 44x73
    // for 4:2:2, 82x73 for 4:4:4), one job per lane and round
    const int cw = a.ss_hor ? SGW : GW, chh = a.ss_ver ? SGH : GH;
    const int cjobs = (cw * chh + kChunk - 1) / kChunk;
    for (int job = t; job < 128 + 2 * cjobs; job += kSetupThreads) {
        const int pl = job < 128 ? 0 : (job < 128 + cjobs ? 1 : 2);
        const int k = pl == 0 ? job : (pl == 1 ? job - 128 : job - 128 - cjobs);
        const int w = pl ? cw : GW, n = pl ? cw * chh : GW * GH;
        const bool on = pl == 0 || (pl == 1 ? want_u : want_v);
        const int first = k * kChunk, last = dmin(first + kChunk, n);
        if (on && first < n) {
            uint32_t s = a.seed;
            if (pl) s ^= pl == 2 ? 0x49d8 : 0xb524;
            s = jump(kJump[k], s & 0xffff);
            int y = first / w, x = first - y * w;
            for (int i = first; i < last; i++) {
                lut[pl][y][x] = (int16_t)round2(g_lds[lfsr(11, s)], shift);
                if (++x == w) {
                    x = 0;
                    y++;
                }
            }
        }
    }
    __syncthreads();
    // the three causal AR filters side by side (chroma after the luma samples
    // it reads), while the fourth wave builds the scaling LUTs and the offsets
    if (wave == 0) {
        ar_wave<false>(lut[0], win[0], (int)d.ar_coeff_shift, gmin, gmax, nullptr, 0, &sm.luma_done, a.err);
    } else if (wave < 3) {
        const int uv = wave - 1;
        const int wl = d.num_points[0] > 0 ? d.ar_coeffs[1 + uv][2 * lag * (lag + 1)] : 0;
        if (uv == 0 ? want_u : want_v) {
            if (a.ss_ver)
                ar_wave<true, 1, 1>(lut[1 + uv], win[1 + uv], (int)d.ar_coeff_shift, gmin, gmax, lut[0], wl,
                                    &sm.luma_done, a.err);
            else if (a.ss_hor)
                ar_wave<true, 1, 0>(lut[1 + uv], win[1 + uv], (int)d.ar_coeff_shift, gmin, gmax, lut[0], wl,
                                    &sm.luma_done, a.err);
            else
                ar_wave<true, 0, 0>(lut[1 + uv], win[1 + uv], (int)d.ar_coeff_shift, gmin, gmax, lut[0], wl,
                                    &sm.luma_done, a.err);
        }
    } else {
        scaling_luts(d, bd, sm.scl, lane);
        band_offsets(a, lane);
    }
    __syncthreads();
    // out: the scaling LUTs and the grain LUTs' sample areas (luma rows 0..72,
    // chroma rows 0..chh-1 at the same row stride), four bytes at a time
    for (int i = t; i < 3 * 4096 / 4; i += kSetupThreads) ((uint32_t *)a.scaling)[i] = ((const uint32_t *)sm.scl)[i];
    constexpr int kLw = GH * GW / 2;
    const int kCw = chh * GW / 2;
    for (int i = t; i < kLw + 2 * kCw; i += kSetupThreads) {
        const int pl = i < kLw ? 0 : (i < kLw + kCw ? 1 : 2);
        const int j = pl * kLw + (pl ? (i - kLw - (pl - 1) * kCw) : i);
        ((uint32_t *)a.lut)[j] = ((const uint32_t *)&lut[0][0][0])[j];
    }
}

__global__ __launch_bounds__(kSetupThreads) void grain_setup_kernel(GrainArgs a, const int16_t *gauss) {
    __shared__ GrainSmem sm;
    grain_setup_body(a, gauss, sm);
}

// Batched set-up: one workgroup per grained frame.
__global__ __launch_bounds__(kSetupThreads) void grain_setup_batch(const GrainArgs *args, const int16_t *gauss) {
    __shared__ GrainSmem sm;
    grain_setup_body(args[blockIdx.x], gauss, sm);
}

// Display samples of one 64x32 tile: one lane = four horizontally adjacent
// samples (a group never straddles a grain block: block sizes are multiples
// of 4), two groups per lane (rows r and r + 16 of the tile).  Every load of
// both groups that does not depend on another load is issued first (the
// lane's word of the scaling LUT, samples, the collocated luma of chroma, the
// block offsets of the block and of its left / upper neighbours), then the
// grain-template reads.  The
// grain sample of block offsets o at (lx, ly) is sample_lut's
// lut[3 + s (3 + o[0]) + ly][3 + s (3 + o[1]) + lx] (s = 2 luma, 1 chroma);
// the overlap rows and columns of a block blend it with the left / upper
// neighbour's (filmgrain.rs:296-380).  The plane's scaling LUT (1 KiB) is
// staged in `scl` (LDS) by one word per lane, so a lookup addresses LDS by the
// sample value itself.
// Words of a plane's scaling LUT each lane stages (1 << BD entries, at least
// one word per lane).
template <int BD> constexpr int scl_words() { return (1 << BD) > 1024 ? (1 << BD) / 1024 : 1; }

// kAny: chroma of any subsampling (a.ss_hor / a.ss_ver); the 4:2:0 / I400
// instantiation has it as a constant.
template <typename P, int BD, bool kAny>
__device__ __forceinline__ void grain_apply_body(const GrainArgs &a, int tile_id, uint8_t *scl) {
    int pl, tx, ty;
    decode_tile(a.map, tile_id, pl, tx, ty);
    const DevPlane &src = a.src[pl], &dst = a.dst[pl];
    const Rav2dFilmGrainData &d = a.fg;
    // one instantiation per bit depth (8-bit frames with P = uint8_t, 10- and
    // 12-bit with uint16_t)
    constexpr int bd = BD, bdm8 = bd - 8, bdmax = (1 << bd) - 1;
    constexpr int ctr = 128 << bdm8, gmin = -ctr, gmax = ctr - 1;
    constexpr int NW = scl_words<BD>();
    // the subsampling of chroma planes (cssx, cssy: constant 1 for 4:2:0) and of this plane
    const int cssx = kAny ? a.ss_hor : 1, cssy = kAny ? a.ss_ver : 1;
    const int ssx = pl ? cssx : 0, ssy = pl ? cssy : 0;
    // log2 of the block width / height in this plane (16 << block_size luma)
    const int lgx = 4 + d.block_size - ssx, lgy = 4 + d.block_size - ssy;
    const int pbsx = 1 << lgx, pbsy = 1 << lgy;
    const int16_t *lut = a.lut + pl * GH * GW;
    const uint8_t *scaling = a.scaling + (pl && d.chroma_scaling_from_luma ? 0 : pl) * 4096;
    const bool on = a.plane_on[pl];
    int minv = 0, maxv = bdmax;
    if (d.clip_to_restricted_range) {
        minv = 16 << bdm8;
        maxv = (pl == 0 || d.mc_identity ? 235 : 240) << bdm8;
    }
    const int W = src.w, H = src.h;
    const uint32_t sstride = (uint32_t)src.stride, dstride = (uint32_t)dst.stride;
    const int x0 = tx * 64, y0 = ty * 32;
    // tile-uniform bases; per-lane offsets are 24-bit products (rows < 2^24 bytes)
    const uint8_t *sbase = src.ptr + (uint64_t)(uint32_t)y0 * sstride + x0 * sizeof(P);
    uint8_t *dbase = dst.ptr + (uint64_t)(uint32_t)y0 * dstride + x0 * sizeof(P);
    const uint32_t lstride = (uint32_t)a.src[0].stride;
    const uint8_t *lbase = a.src[0].ptr + (uint64_t)(uint32_t)(y0 << cssy) * lstride + (x0 << cssx) * sizeof(P);
    const int t = threadIdx.x;
    const int c4 = (t & 15) * 4;  // column of the group in the tile
    // first loads: the scaling LUT words of this lane
    uint32_t sclw[NW];
#pragma unroll
    for (int k = 0; k < NW; k++) sclw[k] = on ? ((const uint32_t *)scaling)[t + 256 * k] : 0u;
    constexpr int NG = 2;
    int gr[NG];
    bool gv[NG];
    uint2 sq[NG], lq0[NG], lq1[NG];
    uint32_t o00[NG], o10[NG], o01[NG], o11[NG];  // packed (y, x) offset bytes of the block, left, upper, upper-left
#pragma unroll
    for (int u = 0; u < NG; u++) {
        const int r = (t >> 4) + 16 * u;
        gr[u] = r;
        gv[u] = x0 + c4 < W && y0 + r < H;
        const int rc = dmin(r, H - 1 - y0), cc = dmin(c4, W - 4 - x0);  // clamped: always valid addresses
        sq[u] = load4((const P *)(sbase + __umul24(rc, sstride)) + cc);
        if (pl) {
            const P *lrow = (const P *)(lbase + __umul24(rc << cssy, lstride)) + (cc << cssx);
            lq0[u] = load4(lrow);
            lq1[u] = cssx ? load4(lrow + 4) : lq0[u];
        }
        if (on) {
            const int band = (y0 + rc) >> lgy, b = (x0 + cc) >> lgx;
            const int bl = dmax(b - 1, 0), bu = dmax(band - 1, 0);
            const uint16_t *ob = (const uint16_t *)a.offsets;
            const int rb = __umul24(band, a.n_blocks), ru = __umul24(bu, a.n_blocks);
            o00[u] = ob[rb + b];
            o10[u] = ob[rb + bl];
            o01[u] = ob[ru + b];
            o11[u] = ob[ru + bl];
        }
    }
    if (!on) {
#pragma unroll
        for (int u = 0; u < NG; u++)
            if (gv[u]) store4((P *)(dbase + __umul24(gr[u], dstride)) + c4, sq[u]);
        return;
    }
#pragma unroll
    for (int k = 0; k < NW; k++) ((uint32_t *)scl)[t + 256 * k] = sclw[k];  // on is tile-uniform: every lane reaches the barrier
    __syncthreads();
    // overlap weights of a direction with / without subsampling (filmgrain.rs:248-592)
    auto Wt = [&](int ss, int k, int j) -> int { return ss ? (k == 0 ? (j == 0 ? 23 : 22) : 0) : ((k == j) ? 27 : 17); };
    const int osx = 2 >> ssx, osy = 2 >> ssy;
    // template index of block offsets `o` (x byte high, y byte low) at (lx, ly)
    auto tix = [&](uint32_t o, int lx, int ly) -> int {
        return __umul24(3 + osy * (3 + (int)(o & 0xff)) + ly, GW) + 3 + osx * (3 + (int)(o >> 8)) + lx;
    };
#pragma unroll
    for (int u = 0; u < NG; u++) {
        const int x = x0 + c4, y = y0 + gr[u];
        int sv[4] = {(int)(sq[u].x & 0xffff), (int)(sq[u].x >> 16), (int)(sq[u].y & 0xffff), (int)(sq[u].y >> 16)};
        int val[4];
        if (pl) {
            const int ls[8] = {(int)(lq0[u].x & 0xffff), (int)(lq0[u].x >> 16), (int)(lq0[u].y & 0xffff),
                               (int)(lq0[u].y >> 16),    (int)(lq1[u].x & 0xffff), (int)(lq1[u].x >> 16),
                               (int)(lq1[u].y & 0xffff), (int)(lq1[u].y >> 16)};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int avg = cssx ? (ls[2 * k] + ls[2 * k + 1] + 1) >> 1 : ls[k];
                if (!d.chroma_scaling_from_luma) {
                    const int uv = pl - 1;
                    const int combined = __mul24(avg, d.uv_luma_mult[uv]) + __mul24(sv[k], d.uv_mult[uv]);
                    val[k] = iclip((combined >> 6) + d.uv_offset[uv] * (1 << bdm8), 0, bdmax);
                } else {
                    val[k] = avg;
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) val[k] = sv[k];
        }
        // samples above the coded maximum (wrapped CDEF results left unrestored)
        // would index past the scaling LUT, where the reference aborts; clamp
        int sc[4];
#pragma unroll
        for (int k = 0; k < 4; k++) sc[k] = scl[dmin(val[k], bdmax)];
        const int band = y >> lgy, b = x >> lgx;
        const int ly = y & (pbsy - 1), lx0 = x & (pbsx - 1);
        const int16_t *g0 = lut + tix(o00[u], lx0, ly);
        int g[4] = {g0[0], g0[1], g0[2], g0[3]};
        // overlap with the left neighbour block (its first osx <= 2 columns)
        // and/or the upper one (first osy rows).  The reference clips the
        // overlap to the block's size at the plane's edge (min(os, bw)); plane
        // sizes are multiples of 4 (frames are multiples of 8), so a clipped
        // block still spans at least 4 > os samples.  lx0 is a multiple of 4:
        // only a group that starts a block blends horizontally, and only its
        // samples 0 and 1 (sample 1 without horizontal subsampling, osx = 2).
        const bool hov = d.overlap_flag && b > 0 && lx0 == 0;
        const bool vov = d.overlap_flag && band > 0 && ly < osy;
        if (hov || vov) {
            auto blend = [&](int n, int c, int wn, int wc) {
                return iclip(round2(__mul24(n, wn) + __mul24(c, wc), 5), gmin, gmax);
            };
            auto hblend = [&](const int16_t *n, int (&c)[4]) {
                c[0] = blend(n[0], c[0], Wt(ssx, 0, 0), Wt(ssx, 0, 1));
                if (!ssx) c[1] = blend(n[1], c[1], Wt(ssx, 1, 0), Wt(ssx, 1, 1));
            };
            if (hov) hblend(lut + tix(o10[u], pbsx, ly), g);
            if (vov) {
                const int16_t *gt = lut + tix(o01[u], lx0, ly + pbsy);
                int top[4] = {gt[0], gt[1], gt[2], gt[3]};
                if (hov) hblend(lut + tix(o11[u], pbsx, ly + pbsy), top);
                const int w0 = Wt(ssy, ly, 0), w1 = Wt(ssy, ly, 1);
#pragma unroll
                for (int k = 0; k < 4; k++) g[k] = blend(top[k], g[k], w0, w1);
            }
        }
        int out[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int noise = round2(__mul24(sc[k], g[k]), d.scaling_shift);  // 8-bit x 12-bit
            out[k] = dmin(dmax(sv[k] + noise, minv), maxv);
        }
        if (gv[u])
            store4((P *)(dbase + __umul24(gr[u], dstride)) + c4,
                   make_uint2((uint32_t)out[0] | ((uint32_t)out[1] << 16), (uint32_t)out[2] | ((uint32_t)out[3] << 16)));
    }
}

template <typename P, int BD, bool kAny> __global__ __launch_bounds__(256) void grain_apply_kernel(GrainArgs a) {
    __shared__ __align__(4) uint8_t scl[1024 * scl_words<BD>()];
    grain_apply_body<P, BD, kAny>(a, blockIdx.x, scl);
}

template <typename P, int BD, bool kLoop, bool kAny>
__global__ __launch_bounds__(256) void grain_apply_batch(const GrainArgs *__restrict__ args, BatchSched s) {
    __shared__ __align__(4) uint8_t scl[1024 * scl_words<BD>()];
    for_each_item<kLoop>(s, [&](int f, int t) {
        grain_apply_body<P, BD, kAny>(args[f], t, scl);
        if (kLoop) __syncthreads();  // the next item restages scl
    });
}

// Per device: the Gaussian sequence in device memory and the LFSR jump-ahead
// matrices in constant memory.
// The Gaussian sequence lives in a static device array (each device's code
// object holds one), filled once per device: building the tables allocates
// nothing.
__device__ __align__(16) int16_t g_gauss[2048];

static const int16_t *grain_tables() {
    static std::mutex mu;
    static const int16_t *per_dev[64] = {};
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    const int16_t *&gauss = per_dev[dev];
    if (!gauss) {
        int16_t *g = nullptr;
        static_assert(sizeof(R2T_GAUSSIAN_SEQUENCE) == sizeof(int16_t) * 2048, "Gaussian sequence size");
        if (hipGetSymbolAddress((void **)&g, HIP_SYMBOL(g_gauss)) != hipSuccess) return nullptr;
        if (hipMemcpy(g, R2T_GAUSSIAN_SEQUENCE, sizeof(R2T_GAUSSIAN_SEQUENCE), hipMemcpyHostToDevice) != hipSuccess)
            return nullptr;
        // jump-ahead matrices: M^(k kChunk), k < 128, and M^(8 kBandSeg j), j < kSegs
        auto mul = [](const uint16_t *A, const uint16_t *B, uint16_t *C) {
            for (int i = 0; i < 16; i++) {
                uint16_t r = 0;
                for (int j = 0; j < 16; j++)
                    if (A[i] >> j & 1) r ^= B[j];
                C[i] = r;
            }
        };
        uint16_t step[16];
        for (int i = 0; i < 15; i++) step[i] = (uint16_t)(1u << (i + 1));
        step[15] = (uint16_t)((1u << 0) | (1u << 1) | (1u << 3) | (1u << 12));
        auto powers = [&](int n, int count, uint16_t (*out)[16]) {  // out[k] = M^(k n)
            uint16_t stepN[16], cur[16], tmp[16];
            for (int i = 0; i < 16; i++) stepN[i] = cur[i] = (uint16_t)(1u << i);
            for (int k = 0; k < n; k++) {
                mul(step, stepN, tmp);
                ::memcpy(stepN, tmp, sizeof(tmp));
            }
            for (int k = 0; k < count; k++) {
                ::memcpy(out[k], cur, sizeof(cur));
                mul(stepN, cur, tmp);
                ::memcpy(cur, tmp, sizeof(tmp));
            }
        };
        static uint16_t jumps[128][16], segs[kSegs][16];
        powers(kChunk, 128, jumps);
        powers(8 * kBandSeg, kSegs, segs);
        (void)hipMemcpyToSymbol(HIP_SYMBOL(kJump), jumps, sizeof(jumps));
        (void)hipMemcpyToSymbol(HIP_SYMBOL(kJumpSeg), segs, sizeof(segs));
        gauss = g;
    }
    return gauss;
}

// generate_grain_y / generate_grain_uv (filmgrain.rs:115-229) of one plane
// into a caller's template buffer (rows of 82): the set-up's Gaussian draws
// and wavefront AR filter for that plane alone; a chroma plane (subsampling
// ss_hor, ss_ver) reads the caller's luma template, complete before the
// filter starts.
struct GrainGenArgs {
    Rav2dFilmGrainData fg;
    uint32_t seed;
    int32_t bitdepth, plane, ss_hor, ss_ver;
    int16_t *out;
    const int16_t *luma;
};

__global__ __launch_bounds__(kSetupThreads) void grain_generate_kernel(GrainGenArgs a, const int16_t *gauss) {
    __shared__ GrainSmem sm;
    const Rav2dFilmGrainData &d = a.fg;
    const int bdm8 = a.bitdepth - 8;
    const int shift = 4 - bdm8 + d.grain_scale_shift;
    const int ctr = 128 << bdm8, gmin = -ctr, gmax = ctr - 1;
    const int t = threadIdx.x, wave = t >> 6, lag = d.ar_coeff_lag, pl = a.plane;
    for (int i = t; i < 2048; i += kSetupThreads) sm.g_lds[i] = gauss[i];
    if (t < 28) {
        const int r = t / 7, c = t % 7, dy = r - 3, dx = c - 3;
        const bool in = dy >= -lag && dx >= -lag && (dy < 0 ? dx <= lag : dx < 0);
        sm.win[pl][r][c] = (int8_t)(in ? d.ar_coeffs[pl][(dy + lag) * (2 * lag + 1) + (dx + lag)] : 0);
    }
    if (pl)
        for (int i = t; i < GH * GW; i += kSetupThreads) sm.lut[0][i / GW][i % GW] = a.luma[i];
    if (t == 0) sm.luma_done = 1 << 30;  // the luma template is complete
    __syncthreads();
    // the plane's template: 73 x 82 luma; chroma 38 x 44 (4:2:0), 73 x 44 (4:2:2), 73 x 82 (4:4:4)
    const int rows = pl && a.ss_ver ? SGH : GH, cols = pl && a.ss_hor ? SGW : GW;
    {
        const int w = cols, n = cols * rows, lanes = (n + kChunk - 1) / kChunk;  // at most 128 (kJump)
        const int first = t * kChunk, last = dmin(first + kChunk, n);
        if (t < lanes) {
            uint32_t s = a.seed;
            if (pl) s ^= pl == 2 ? 0x49d8 : 0xb524;
            s = jump(kJump[t], s & 0xffff);
            int y = first / w, x = first - y * w;
            for (int i = first; i < last; i++) {
                sm.lut[pl][y][x] = (int16_t)round2(sm.g_lds[lfsr(11, s)], shift);
                if (++x == w) {
                    x = 0;
                    y++;
                }
            }
        }
    }
    __syncthreads();
    if (wave == 0) {
        if (pl == 0) {
            ar_wave<false>(sm.lut[0], sm.win[0], (int)d.ar_coeff_shift, gmin, gmax, nullptr, 0, &sm.luma_done, nullptr);
        } else {
            const int wl = d.num_points[0] > 0 ? d.ar_coeffs[pl][2 * lag * (lag + 1)] : 0;
            if (a.ss_ver)
                ar_wave<true, 1, 1>(sm.lut[pl], sm.win[pl], (int)d.ar_coeff_shift, gmin, gmax, sm.lut[0], wl,
                                    &sm.luma_done, nullptr);
            else if (a.ss_hor)
                ar_wave<true, 1, 0>(sm.lut[pl], sm.win[pl], (int)d.ar_coeff_shift, gmin, gmax, sm.lut[0], wl,
                                    &sm.luma_done, nullptr);
            else
                ar_wave<true, 0, 0>(sm.lut[pl], sm.win[pl], (int)d.ar_coeff_shift, gmin, gmax, sm.lut[0], wl,
                                    &sm.luma_done, nullptr);
        }
    }
    __syncthreads();
    for (int i = t; i < rows * cols; i += kSetupThreads) {
        const int y = i / cols, x = i - y * cols;
        a.out[y * GW + x] = sm.lut[pl][y][x];
    }
}

int launch_grain_generate(const Rav2dFilmGrainData &fg, uint32_t seed, int bitdepth, int plane, int ss_hor, int ss_ver,
                          int16_t *out, const int16_t *luma, hipStream_t s) {
    const int16_t *gauss = grain_tables();
    if (!gauss) return RAV2D_ENOMEM;
    GrainGenArgs a;
    a.fg = fg;
    a.seed = seed;
    a.bitdepth = bitdepth;
    a.plane = plane;
    a.ss_hor = ss_hor;
    a.ss_ver = ss_ver;
    a.out = out;
    a.luma = luma;
    hipLaunchKernelGGL(grain_generate_kernel, dim3(1), dim3(kSetupThreads), 0, s, a, gauss);
    return hipGetLastError() == hipSuccess ? RAV2D_OK : RAV2D_EDEVICE;
}

void launch_grain_setup(const GrainArgs &a, hipStream_t s) {
    const int16_t *gauss = grain_tables();
    if (!gauss) return;
    R2G_LAUNCH(grain_setup_kernel, dim3(1), dim3(kSetupThreads), 0, s, a, gauss);
}

bool grain_init_tables() { return grain_tables() != nullptr; }

void launch_grain_setup_batch(const GrainArgs *d_args, int n, hipStream_t s) {
    const int16_t *gauss = grain_tables();
    if (!gauss || n <= 0) return;
    R2G_LAUNCH(grain_setup_batch, dim3(n), dim3(kSetupThreads), 0, s, d_args, gauss);
}

// Mixed launch: the tiles of every bit-depth group's frames.
template <bool kLoop, bool kAny>
__global__ __launch_bounds__(256) void grain_apply_mix(DevArgs3<GrainArgs> args, MixSched s) {
    __shared__ __align__(4) uint8_t scl[1024 * scl_words<12>()];
    for_each_mixed_item<kLoop>(s, [&](int g, int f, int t) {
        const GrainArgs &fa = args.p[g][f];
        if (g == 0)
            grain_apply_body<uint8_t, 8, kAny>(fa, t, scl);
        else if (g == 1)
            grain_apply_body<uint16_t, 10, kAny>(fa, t, scl);
        else
            grain_apply_body<uint16_t, 12, kAny>(fa, t, scl);
        if (kLoop) __syncthreads();  // the next item restages scl
    });
}

void launch_grain_apply_mix(DevArgs3<GrainArgs> args, const MixSched &s, int grid, bool any, hipStream_t st) {
    if (s.first[3] <= 0) return;
    const bool loop = grid < s.first[3];
    if (loop && any)
        R2G_LAUNCH((grain_apply_mix<true, true>), dim3(grid), dim3(256), 0, st, args, s);
    else if (loop)
        R2G_LAUNCH((grain_apply_mix<true, false>), dim3(grid), dim3(256), 0, st, args, s);
    else if (any)
        R2G_LAUNCH((grain_apply_mix<false, true>), dim3(grid), dim3(256), 0, st, args, s);
    else
        R2G_LAUNCH((grain_apply_mix<false, false>), dim3(grid), dim3(256), 0, st, args, s);
}

void launch_grain_apply_batch(const GrainArgs *d_args, const BatchSched &sc, int grid, int bitdepth, bool any,
                              hipStream_t s) {
    if (sc.total <= 0) return;
    const bool loop = grid < sc.total;
#define R2G_GRAIN_BATCH(P, BD, A)                                                                  \
    do {                                                                                           \
        if (loop)                                                                                  \
            R2G_LAUNCH((grain_apply_batch<P, BD, true, A>), dim3(grid), dim3(256), 0, s, d_args, sc);  \
        else                                                                                       \
            R2G_LAUNCH((grain_apply_batch<P, BD, false, A>), dim3(grid), dim3(256), 0, s, d_args, sc); \
    } while (0)
#define R2G_GRAIN_BATCH_BD(A)              \
    do {                                   \
        if (bitdepth == 12)                \
            R2G_GRAIN_BATCH(uint16_t, 12, A); \
        else if (bitdepth == 10)           \
            R2G_GRAIN_BATCH(uint16_t, 10, A); \
        else                               \
            R2G_GRAIN_BATCH(uint8_t, 8, A);   \
    } while (0)
    if (any)
        R2G_GRAIN_BATCH_BD(true);
    else
        R2G_GRAIN_BATCH_BD(false);
#undef R2G_GRAIN_BATCH_BD
#undef R2G_GRAIN_BATCH
}

void grain_tile_map(GrainArgs &a) {
    a.map.first[0] = 0;
    for (int p = 0; p < 3; p++) {
        a.map.tiles_x[p] = (a.src[p].w + 63) / 64;
        a.map.tiles_y[p] = (a.src[p].h + 31) / 32;
        a.map.first[p + 1] = a.map.first[p] + a.map.tiles_x[p] * a.map.tiles_y[p];
    }
}

void launch_grain_apply(GrainArgs a, int bitdepth, hipStream_t s) {
    grain_tile_map(a);
    const dim3 grid(a.map.first[3]);
#define R2G_GRAIN_ONE(A)                                                                      \
    do {                                                                                      \
        if (bitdepth == 12)                                                                   \
            R2G_LAUNCH((grain_apply_kernel<uint16_t, 12, A>), grid, dim3(256), 0, s, a);      \
        else if (bitdepth == 10)                                                              \
            R2G_LAUNCH((grain_apply_kernel<uint16_t, 10, A>), grid, dim3(256), 0, s, a);      \
        else                                                                                  \
            R2G_LAUNCH((grain_apply_kernel<uint8_t, 8, A>), grid, dim3(256), 0, s, a);        \
    } while (0)
    if (a.ss_hor && a.ss_ver)
        R2G_GRAIN_ONE(false);
    else
        R2G_GRAIN_ONE(true);
#undef R2G_GRAIN_ONE
}

}  // namespace r2g
