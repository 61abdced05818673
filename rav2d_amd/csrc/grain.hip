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

__global__ __launch_bounds__(256) void grain_setup_kernel(GrainArgs a, const int16_t *gauss) {
    __shared__ int16_t lut[3][GH][GW];
    const Rav2dFilmGrainData &d = a.fg;
    const int bd = a.bitdepth, bdm8 = bd - 8;
    const int shift = 4 - bdm8 + d.grain_scale_shift;
    const int ctr = 128 << bdm8, gmin = -ctr, gmax = ctr - 1;
    const int t = threadIdx.x;
    const bool want_u = d.num_points[1] > 0 || d.chroma_scaling_from_luma;
    const bool want_v = d.num_points[2] > 0 || d.chroma_scaling_from_luma;
    // Gaussian draws: one sequential LFSR walk per plane
    if (t < 3) {
        uint32_t s = a.seed;
        int w = GW, h = GH;
        bool on = true;
        if (t > 0) {
            s ^= t == 2 ? 0x49d8 : 0xb524;
            w = SGW;
            h = SGH;
            on = t == 1 ? want_u : want_v;
        }
        for (int y = 0; y < GH; y++)
            for (int x = 0; x < GW; x++) lut[t][y][x] = 0;
        if (on)
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++) lut[t][y][x] = (int16_t)round2(gauss[lfsr(11, s)], shift);
    }
    // scaling LUTs (one lane per plane)
    if (t >= 64 && t < 67) {
        int pl = t - 64;
        uint8_t *sc = a.scaling + pl * 1024;
        int num = d.num_points[pl];
        const uint8_t(*pts)[2] = d.points[pl];
        int size = 1 << bd;
        int sx = bd - 8;
        if (num == 0) {
            for (int i = 0; i < size; i++) sc[i] = 0;
        } else {
            for (int i = 0; i < (pts[0][0] << sx); i++) sc[i] = pts[0][1];
            for (int i = 0; i < num - 1; i++) {
                int bx = pts[i][0], by = pts[i][1], dx = pts[i + 1][0] - bx, dy = pts[i + 1][1] - by;
                if (dx <= 0) continue;
                int delta = dy * ((0x10000 + (dx >> 1)) / dx);
                int dd = 0x8000;
                for (int x = 0; x < dx; x++) {
                    sc[(bx + x) << sx] = (uint8_t)(by + (dd >> 16));
                    dd += delta;
                }
            }
            for (int i = pts[num - 1][0] << sx; i < size; i++) sc[i] = pts[num - 1][1];
            if (sx) {
                int pad = 1 << sx, rnd = pad >> 1;
                for (int i = 0; i < num - 1; i++) {
                    int bx = pts[i][0] << sx, dx = (pts[i + 1][0] << sx) - bx;
                    for (int x = 0; x < dx; x += pad) {
                        int base = sc[bx + x], range = sc[bx + x + pad] - base, r = rnd;
                        for (int n = 1; n < pad; n++) {
                            r += range;
                            sc[bx + x + n] = (uint8_t)(base + (r >> sx));
                        }
                    }
                }
            }
        }
    }
    // block offsets: one lane per row band
    for (int band = t; band < a.n_bands; band += 256) {
        uint32_t s = a.seed;
        s ^= (uint32_t)(((band * 37 + 178) & 0xFF)) << 8;
        s ^= (uint32_t)((band * 173 + 105) & 0xFF);
        uint8_t *o = a.offsets + (size_t)band * a.n_blocks * 2;
        for (int b = 0; b < a.n_blocks; b++)
            for (int n = 0; n < 2; n++) {
                o[b * 2 + n] = (uint8_t)(((uint32_t)(3 - d.block_size) * lfsr(9, s)) >> 6);
                lfsr(16, s);
                lfsr(16, s);
                lfsr(16, s);
            }
    }
    __syncthreads();
    // causal AR filter, luma: sample (y, x) at step x + 4y
    const int lag = d.ar_coeff_lag;
    for (int step = 3 + 4 * 3; step <= (GW - 4) + 4 * (GH - 1); step++) {
        int y = t + 3;
        if (y < GH) {
            int x = step - 4 * y;
            if (x >= 3 && x < GW - 3) {
                int sum = 0, ci = 0;
                for (int dy = y - lag; dy <= y; dy++) {
                    int dx_end = dy == y ? x : x + lag + 1;
                    for (int dx = x - lag; dx < dx_end; dx++) sum += d.ar_coeffs[0][ci++] * lut[0][dy][dx];
                }
                lut[0][y][x] = (int16_t)iclip(lut[0][y][x] + round2(sum, (int)d.ar_coeff_shift), gmin, gmax);
            }
        }
        __syncthreads();
    }
    // chroma (4:2:0 sub-grid), with the luma correlation term
    for (int step = 3 + 4 * 3; step <= (SGW - 4) + 4 * (SGH - 1); step++) {
        int uv = t >> 7, y = (t & 127) + 3;
        bool on = uv == 0 ? want_u : want_v;
        if (on && y < SGH) {
            int x = step - 4 * y;
            if (x >= 3 && x < SGW - 3) {
                const int8_t *c = d.ar_coeffs[1 + uv];
                int sum = 0, ci = 0;
                for (int dy = y - lag; dy <= y; dy++) {
                    int dx_end = dy == y ? x : x + lag + 1;
                    for (int dx = x - lag; dx < dx_end; dx++) sum += c[ci++] * lut[1 + uv][dy][dx];
                }
                if (d.num_points[0] > 0) {
                    int lx = ((x - 3) << 1) + 3, ly = ((y - 3) << 1) + 3;
                    int l = lut[0][ly][lx] + lut[0][ly][lx + 1] + lut[0][ly + 1][lx] + lut[0][ly + 1][lx + 1];
                    sum += round2(l, 2) * c[ci];
                }
                lut[1 + uv][y][x] = (int16_t)iclip(lut[1 + uv][y][x] + round2(sum, (int)d.ar_coeff_shift), gmin, gmax);
            }
        }
        __syncthreads();
    }
    for (int i = t; i < 3 * GH * GW; i += 256) a.lut[i] = (&lut[0][0][0])[i];
}

// One output sample: S(bxo, byo) samples the block `bxo` to the left and
// `byo` above, as sample_lut with offsets[bxo][byo].
template <typename P>
__global__ __launch_bounds__(256) void grain_apply_kernel(GrainArgs a) {
    int pl, tx, ty;
    decode_tile(a.map, blockIdx.x, pl, tx, ty);
    const DevPlane &src = a.src[pl], &dst = a.dst[pl];
    const Rav2dFilmGrainData &d = a.fg;
    const int bd = a.bitdepth, bdm8 = bd - 8;
    const int ctr = 128 << bdm8, gmin = -ctr, gmax = ctr - 1;
    const int bs = 16 << d.block_size;
    const int sub = pl ? 1 : 0;
    const int pbs = bs >> sub;  // block size in this plane
    const int16_t(*lut)[GW] = (const int16_t(*)[GW])(a.lut + pl * GH * GW);
    const uint8_t *scaling = a.scaling + (pl && d.chroma_scaling_from_luma ? 0 : pl) * 1024;
    const bool on = a.plane_on[pl];
    int minv = 0, maxv = (1 << bd) - 1;
    if (d.clip_to_restricted_range) {
        minv = 16 << bdm8;
        maxv = (pl == 0 || d.mc_identity ? 235 : 240) << bdm8;
    }
    for (int i = threadIdx.x; i < 64 * 32; i += 256) {
        int x = tx * 64 + (i & 63), y = ty * 32 + (i >> 6);
        if (x >= src.w || y >= src.h) continue;
        int sv = row_ptr<const P>(src, y)[x];
        if (!on) {
            row_ptr<P>(dst, y)[x] = (P)sv;
            continue;
        }
        int band = y / pbs, b = x / pbs;
        int ly = y - band * pbs, lx = x - b * pbs;
        int bh = dmin(pbs, src.h - band * pbs), bw = dmin(pbs, src.w - b * pbs);
        int ystart = (d.overlap_flag && band > 0) ? dmin(2 >> sub, bh) : 0;
        int xstart = (d.overlap_flag && b > 0) ? dmin(2 >> sub, bw) : 0;
        auto S = [&](int bxo, int byo) -> int {
            const uint8_t *o = a.offsets + ((size_t)(band - byo) * a.n_blocks + (b - bxo)) * 2;
            int offx = 3 + (2 >> sub) * (3 + o[1]);
            int offy = 3 + (2 >> sub) * (3 + o[0]);
            return lut[offy + ly + pbs * byo][offx + lx + pbs * bxo];
        };
        // overlap weights: [27,17],[17,27] (full-size); [23,22] (subsampled)
        auto W = [&](int k, int j) -> int { return sub ? (k == 0 ? (j == 0 ? 23 : 22) : 0) : ((k == j) ? 27 : 17); };
        int g = S(0, 0);
        if (lx < xstart) g = iclip(round2(S(1, 0) * W(lx, 0) + g * W(lx, 1), 5), gmin, gmax);
        if (ly < ystart) {
            int top = S(0, 1);
            if (lx < xstart) top = iclip(round2(S(1, 1) * W(lx, 0) + top * W(lx, 1), 5), gmin, gmax);
            g = iclip(round2(top * W(ly, 0) + g * W(ly, 1), 5), gmin, gmax);
        }
        int val = sv;
        if (pl) {
            const int16_t *unused = nullptr;
            (void)unused;
            const P *lrow = row_ptr<const P>(a.src[0], 2 * y);
            int l = lrow[2 * x];
            int avg = (l + lrow[2 * x + 1] + 1) >> 1;
            if (!d.chroma_scaling_from_luma) {
                int uv = pl - 1;
                int combined = avg * d.uv_luma_mult[uv] + sv * d.uv_mult[uv];
                val = iclip((combined >> 6) + d.uv_offset[uv] * (1 << bdm8), 0, (1 << bd) - 1);
            } else {
                val = avg;
            }
        }
        // samples above the coded maximum (wrapped CDEF results left unrestored)
        // would index past the scaling LUT, where the reference aborts; clamp
        int noise = round2(scaling[dmin(val, (1 << bd) - 1)] * g, d.scaling_shift);
        row_ptr<P>(dst, y)[x] = (P)iclip(sv + noise, minv, maxv);
    }
}

void launch_grain(GrainArgs a, int bitdepth, hipStream_t s) {
    static const int16_t *gauss = nullptr;
    if (!gauss) {
        int16_t *g = nullptr;
        if (hipMalloc((void **)&g, sizeof(R2T_GAUSSIAN_SEQUENCE)) != hipSuccess) return;
        (void)hipMemcpy(g, R2T_GAUSSIAN_SEQUENCE, sizeof(R2T_GAUSSIAN_SEQUENCE), hipMemcpyHostToDevice);
        gauss = g;
    }
    hipLaunchKernelGGL(grain_setup_kernel, dim3(1), dim3(256), 0, s, a, gauss);
    a.map.first[0] = 0;
    for (int p = 0; p < 3; p++) {
        a.map.tiles_x[p] = (a.src[p].w + 63) / 64;
        a.map.tiles_y[p] = (a.src[p].h + 31) / 32;
        a.map.first[p + 1] = a.map.first[p] + a.map.tiles_x[p] * a.map.tiles_y[p];
    }
    if (bitdepth > 8)
        hipLaunchKernelGGL(grain_apply_kernel<uint16_t>, dim3(a.map.first[3]), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(grain_apply_kernel<uint8_t>, dim3(a.map.first[3]), dim3(256), 0, s, a);
}

}  // namespace r2g
