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
#include "internal.h"
#include "tables.h"

namespace r2g {

__constant__ int8_t kNsY[16][2] = {{1, 0}, {0, 1}, {2, 0}, {0, 2}, {1, 1}, {-1, 1}, {2, 1}, {2, -1},
                                   {1, 2}, {1, -2}, {3, 0}, {0, 3}, {4, 0}, {0, 4}, {3, 3}, {3, -3}};
__constant__ int8_t kPc[12][2] = {{1, 0}, {0, 1}, {2, 0}, {0, 2}, {1, 1}, {-1, 1},
                                  {2, 1}, {2, -1}, {1, 2}, {1, -2}, {3, 0}, {0, 3}};
__constant__ int8_t kNsUv[6][2] = {{1, 0}, {0, 1}, {1, 1}, {-1, 1}, {2, 0}, {0, 2}};
__constant__ int8_t kNsUvY[12][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, -1},
                                     {-1, 1}, {1, -1}, {2, 0}, {-2, 0}, {0, 2}, {0, -2}};
__constant__ int8_t kGdfCoords[18][2] = {{6, 0}, {5, 0}, {4, 0}, {3, 0}, {2, 1}, {2, 0}, {2, -1}, {1, 2}, {1, 1},
                                         {1, 0}, {1, -1}, {1, -2}, {0, 6}, {0, 5}, {0, 4}, {0, 3}, {0, 2}, {0, 1}};
__constant__ uint16_t kPcNorm[4] = {3739, 3273, 3074, 7};
__constant__ int16_t kModeW[4][3] = {{-527, 15325, 321}, {26436, -17705, 17905}, {366, -147, -194}, {202, -267, -179}};
__constant__ int16_t kModeOff[4] = {-547, -21565, -573, -680};

constexpr int kThreads = 256;
constexpr int E = 6;               // staged halo (GDF reach)
constexpr int VW = 64 + 2 * E;     // staged width
constexpr int VH = 64 + 2 * E;     // staged height

__device__ __forceinline__ int qval_tskip(int qstep, int tskip, int i, int bdm8) {
    qstep = (qstep + ((1 << bdm8) >> 1)) >> bdm8;
    int prod = (tskip * qstep + 128) >> 8;
    int q = kModeW[i][0] * (tskip << 5) + kModeW[i][1] * qstep + kModeW[i][2] * prod;
    q = apply_sign((abs(q) + (1 << 12)) >> 13, q);
    return 255 * (kModeOff[i] + q);
}

struct LrSmem {
    uint16_t V[VH * VW];
    uint16_t Ld[(32 + 4) * (64 + 4)];
    uint16_t cells[33 * 33 * 4];
    uint8_t cls[16 * 16];
    uint16_t s_alpha[88];
    int16_t s_weight[264];
    LrItem it;
};

template <typename P>
__device__ void lr_body(const LrArgs &a, int item, LrSmem &sm) {
    uint16_t *V = sm.V, *Ld = sm.Ld, *cells = sm.cells, *s_alpha = sm.s_alpha;
    uint8_t *cls = sm.cls;
    int16_t *s_weight = sm.s_weight;
    LrItem &it = sm.it;
    if (threadIdx.x < sizeof(LrItem) / 4) ((uint32_t *)&it)[threadIdx.x] = ((const uint32_t *)&a.items[item])[threadIdx.x];
    __syncthreads();
    const int pl = it.plane, x0 = it.x0, w = it.w, y0 = it.y0, h = it.h;
    const uint8_t edges = it.edges;
    const DevPlane &f3 = a.f3[pl], &f2 = a.f2[pl], &dst = a.dst[pl];
    const int bd = a.g.bitdepth, bdm8 = bd - 8, bdmax = a.g.bdmax;
    const int kind = it.kind;
    const bool gdf = it.gdf_on;
    const int VS = w + 2 * E;  // row stride of V
    if (kind == LR_KIND_NONE && !gdf) {
        for (int i = threadIdx.x; i < w * h; i += kThreads) {
            int y = i / w, x = i - y * w;
            row_ptr<P>(dst, y0 + y)[x0 + x] = row_ptr<const P>(f3, y0 + y)[x0 + x];
        }
        return;
    }
    // ---- stage the virtual image V: rows [-E, h+E), columns [-E, w+E)
    for (int i = threadIdx.x; i < (h + 2 * E) * VS; i += kThreads) {
        int r = i / VS - E, c = i % VS - E;
        int xs = (c < 0 && !(edges & LRE_LEFT)) ? 0 : ((c >= w && !(edges & LRE_RIGHT)) ? w - 1 : c);
        int gx = x0 + xs;
        P v;
        if (r >= 0 && r < h)
            v = row_ptr<const P>(f3, y0 + r)[gx];
        else if (r < 0)
            v = (edges & LRE_TOP) ? row_ptr<const P>(f2, y0 + (r == -1 ? -1 : -2))[gx] : row_ptr<const P>(f3, y0)[gx];
        else
            v = (edges & LRE_BOTTOM) ? row_ptr<const P>(f2, y0 + h + (r == h ? 0 : 1))[gx]
                                     : row_ptr<const P>(f3, y0 + h - 1)[gx];
        V[i] = (uint16_t)v;
    }
    const bool uv_single = kind == LR_KIND_SINGLE && pl != 0;
    const int LS = w + 4;
    if (uv_single) {
        // downsampled luma rows of the chroma stripe (cfl_ds_filter_index)
        const DevPlane &l3 = a.f3[0], &l2 = a.f2[0];
        for (int i = threadIdx.x; i < (h + 4) * LS; i += kThreads) {
            int r = i / LS - 2, c = i % LS - 2;
            int xs = (c < 0 && !(edges & LRE_LEFT)) ? 0 : ((c >= w && !(edges & LRE_RIGHT)) ? w - 1 : c);
            int lx = 2 * (x0 + xs);
            const P *ra, *rb;
            int rr = r;
            if (rr >= h && !(edges & LRE_BOTTOM)) rr = h - 1;
            if (rr < 0 && !(edges & LRE_TOP)) rr = 0;
            if (rr >= 0 && rr < h) {
                ra = row_ptr<const P>(l3, 2 * (y0 + rr));
                rb = row_ptr<const P>(l3, 2 * (y0 + rr) + 1);
            } else if (rr == -2) {
                ra = rb = row_ptr<const P>(l2, 2 * y0 - 2);
            } else if (rr == -1) {
                ra = row_ptr<const P>(l2, 2 * y0 - 2);
                rb = row_ptr<const P>(l2, 2 * y0 - 1);
            } else if (rr == h) {
                ra = row_ptr<const P>(l2, 2 * (y0 + h));
                rb = row_ptr<const P>(l2, 2 * (y0 + h) + 1);
            } else {
                ra = rb = row_ptr<const P>(l2, 2 * (y0 + h) + 1);
            }
            int v;
            if (a.cfl_ds == 0)
                v = ((int)ra[lx] + rb[lx] + ra[lx + 1] + rb[lx + 1]) >> 2;
            else if (a.cfl_ds == 1)
                v = ((int)ra[lx] + rb[lx]) >> 1;
            else
                v = ra[lx];
            Ld[i] = (uint16_t)v;
        }
    }
    if (gdf) {
        int ab = a.gdf_ref * 528 + a.gdf_qp * 88, wb = a.gdf_ref * 1584 + a.gdf_qp * 264;
        for (int i = threadIdx.x; i < 88; i += kThreads) s_alpha[i] = a.t.gdf_alpha[ab + i];
        for (int i = threadIdx.x; i < 264; i += kThreads) s_weight[i] = a.t.gdf_weight[wb + i];
    }
    __syncthreads();
    auto VA = [&](int r, int c) -> int { return V[(r + E) * VS + c + E]; };
    const int down = bd == 12 ? 2 : 0, up = bd == 8 ? 2 : 0;
    // ---- GDF gradient cells: cell (j,k) covers centre columns 2j-1,2j and rows 2k-1,2k
    const int ncx = w / 2 + 1, ncy = h / 2 + 1;
    if (gdf) {
        for (int i = threadIdx.x; i < ncx * ncy; i += kThreads) {
            int k = i / ncx, j = i - k * ncx;
            int g[4] = {0, 0, 0, 0};
#pragma unroll
            for (int yy = 0; yy < 2; yy++)
#pragma unroll
                for (int xx = 0; xx < 2; xx++) {
                    int cy = 2 * k - 1 + yy, cx = 2 * j - 1 + xx;
                    int b = VA(cy, cx) >> down;
                    g[0] += abs(b * 2 - (VA(cy - 1, cx) >> down) - (VA(cy + 1, cx) >> down));
                    g[1] += abs(b * 2 - (VA(cy, cx - 1) >> down) - (VA(cy, cx + 1) >> down));
                    g[2] += abs(b * 2 - (VA(cy - 1, cx - 1) >> down) - (VA(cy + 1, cx + 1) >> down));
                    g[3] += abs(b * 2 - (VA(cy + 1, cx - 1) >> down) - (VA(cy - 1, cx + 1) >> down));
                }
#pragma unroll
            for (int d = 0; d < 4; d++) cells[i * 4 + d] = (uint16_t)g[d];
        }
    }
    // ---- classes of the multi-class filters (one per 4x4 block)
    const int bw4 = w >> 2, bh4 = h >> 2;
    if (kind == LR_KIND_MULTI || kind == LR_KIND_PC) {
        const uint8_t *sub = kind == LR_KIND_PC ? a.t.pc_sub + a.widx * 256 : a.t.ns_sub + (a.widx * 7 + a.ns_idx) * 256;
        for (int i = threadIdx.x; i < bw4 * bh4; i += kThreads) {
            int by = i / bw4, bx = i - by * bw4;
            int f0 = 0, f1 = 0, f2v = 0, s = 0;
            for (int dy = -1; dy <= 4; dy++)
                for (int dx = -1; dx <= 4; dx++) {
                    int x = bx * 4 + dx, y = by * 4 + dy;
                    int m = VA(y, x);
                    f0 += abs(VA(y - 1, x) - 2 * m + VA(y + 1, x));
                    f1 += abs(VA(y - 1, x + 1) - 2 * m + VA(y + 1, x - 1));
                    f2v += abs(VA(y - 1, x - 1) - 2 * m + VA(y + 1, x + 1));
                }
            const int npx[3] = {16, 4, 1};
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int e = (dy != 0) + (dx != 0);
                    int fx = iclip((bx & 15) + dx, 0, 15), fy = iclip(by + dy, 0, bh4 - 1);
                    s += npx[e] * (((it.noskip[fy] >> fx) & 1) == 0);
                }
            int rnd = (1 << bdm8) >> 1;
            int f[3] = {f0, f1, f2v};
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
            cls[by * 16 + bx] = sub[a.t.lut_to_class[li]];
        }
    }
    __syncthreads();
    // ---- outputs
    const int sh = 12 - bd, rnd_g = (1 << sh) >> 1;
    int err_base, gscale;
    if (a.gdf_ref == 0) {
        err_base = a.gdf_qp * 4096;
        gscale = 8;
    } else {
        err_base = (a.gdf_ref - 1) * 6000 + a.gdf_qp * 1000;
        gscale = 5;
    }
    const int bias = a.gdf_ref * 6 + a.gdf_qp;
    for (int i = threadIdx.x; i < w * h; i += kThreads) {
        int y = i / w, x = i - y * w;
        bool lossless = (it.ll[y >> 2] >> (x >> 2)) & 1;
        int m = VA(y, x);
        int out = m;
        if (!lossless) {
            if (kind == LR_KIND_SINGLE && pl == 0) {
                int s = m << 7;
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    int dy = kNsY[k][0], dx = kNsY[k][1];
                    s += (VA(y + dy, x + dx) + VA(y - dy, x - dx) - 2 * m) * it.filt[k];
                }
                out = iclip((s + 64) >> 7, 0, bdmax);
            } else if (kind == LR_KIND_SINGLE) {
                int s = m << 7;
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    int dy = kNsUv[k][0], dx = kNsUv[k][1];
                    s += (VA(y + dy, x + dx) + VA(y - dy, x - dx) - 2 * m) * it.filt[k];
                }
                int l = Ld[(y + 2) * LS + x + 2];
#pragma unroll
                for (int k = 0; k < 12; k++) {
                    int dy = kNsUvY[k][0], dx = kNsUvY[k][1];
                    s += ((int)Ld[(y + 2 + dy) * LS + x + 2 + dx] - l) * it.filt[6 + k];
                }
                out = iclip((s + 64) >> 7, 0, bdmax);
            } else if (kind == LR_KIND_MULTI) {
                const int8_t *fl = a.mfilt[pl][cls[(y >> 2) * 16 + (x >> 2)]];
                int s = m << 7;
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    int dy = kNsY[k][0], dx = kNsY[k][1];
                    s += (VA(y + dy, x + dx) + VA(y - dy, x - dx) - 2 * m) * fl[k];
                }
                out = iclip((s + 64) >> 7, 0, bdmax);
            } else if (kind == LR_KIND_PC) {
                const int16_t *fl = a.t.pc_filters + (a.widx * 64 + cls[(y >> 2) * 16 + (x >> 2)]) * 13;
                int s = m * fl[12];
#pragma unroll
                for (int k = 0; k < 12; k++) {
                    int dy = kPc[k][0], dx = kPc[k][1];
                    s += (VA(y + dy, x + dx) + VA(y - dy, x - dx)) * fl[k];
                }
                out = iclip((s + 64) >> 7, 0, bdmax);
            }
            if (gdf) {
                int x1 = x & ~1, hx = x1 >> 1, ky = (y & ~1) >> 1;
                int gs[4];
#pragma unroll
                for (int d = 0; d < 4; d++)
                    gs[d] = cells[(ky * ncx + hx) * 4 + d] + cells[(ky * ncx + hx + 1) * 4 + d] +
                            cells[((ky + 1) * ncx + hx) * 4 + d] + cells[((ky + 1) * ncx + hx + 1) * 4 + d];
                int c = (gs[0] <= gs[1]) | ((gs[2] <= gs[3]) << 1);
                int iv[3];
#pragma unroll
                for (int k = 0; k < 3; k++) iv[k] = a.t.gdf_bias[bias * 3 + k];
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    int kk = d + 18;
                    int alpha = s_alpha[kk * 4 + c];
                    int v = dmin(gs[d] >> (4 - up), alpha);
#pragma unroll
                    for (int k = 0; k < 3; k++) iv[k] += v * s_weight[k * 88 + kk * 4 + c];
                }
                int mm = m >> down;
#pragma unroll
                for (int kk = 0; kk < 18; kk++) {
                    int alpha = s_alpha[kk * 4 + c];
                    int dy = kGdfCoords[kk][0], dx = kGdfCoords[kk][1];
                    int av = VA(y - dy, x - dx) >> down, bv = VA(y + dy, x + dx) >> down;
                    int above = iclip((av - mm) * (1 << up), -alpha, alpha);
                    int below = iclip((bv - mm) * (1 << up), -alpha, alpha);
                    int v = iclip(above + below, -512, 511);
#pragma unroll
                    for (int k = 0; k < 3; k++) iv[k] += v * s_weight[k * 88 + kk * 4 + c];
                }
                int full = 0;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    int sv = iv[k] * gscale;
                    int v = apply_sign((abs(sv) + (1 << 14)) >> 15, sv);
                    full = full * (gscale * 2) + iclip(v, -gscale, gscale - 1) + gscale;
                }
                int err = a.gdf_ref == 0 ? a.t.gdf_intra[err_base + full] : a.t.gdf_inter[err_base + full];
                int diff = err * a.gdf_scale;
                out = iclip(out + apply_sign((abs(diff) + rnd_g) >> sh, diff), 0, bdmax);
            }
        }
        row_ptr<P>(dst, y0 + y)[x0 + x] = (P)out;
    }
}

template <typename P> __global__ __launch_bounds__(kThreads) void lr_kernel(LrArgs a) {
    __shared__ LrSmem sm;
    lr_body<P>(a, blockIdx.x, sm);
}

// Batched launch: every frame's items (restoration units and plane copies).
__global__ __launch_bounds__(kThreads) void lr_batch(const LrArgs *args, BatchSched s) {
    __shared__ LrSmem sm;
    for_each_item(s, [&](int f, int t) {
        const LrArgs &a = args[f];
        if (a.g.bitdepth > 8)
            lr_body<uint16_t>(a, t, sm);
        else
            lr_body<uint8_t>(a, t, sm);
        __syncthreads();
    });
}

void launch_lr_batch(const LrArgs *d_args, const BatchSched &sc, int grid, hipStream_t s) {
    if (sc.total <= 0) return;
    hipLaunchKernelGGL(lr_batch, dim3(grid), dim3(kThreads), 0, s, d_args, sc);
}

void launch_lr(const LrArgs &a, int n_items, int bitdepth, hipStream_t s) {
    if (n_items <= 0) return;
    if (bitdepth > 8)
        hipLaunchKernelGGL(lr_kernel<uint16_t>, dim3(n_items), dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL(lr_kernel<uint8_t>, dim3(n_items), dim3(kThreads), 0, s, a);
}

// Device copies of the tables, made once per device.
int lr_tables(int device, LrTables *out) {
    static LrTables cache[16];
    static bool ready[16] = {false};
    if (device < 0 || device >= 16) return RAV2D_EINVAL;
    if (!ready[device]) {
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
        };
        size_t total = 0, off[9];
        for (int i = 0; i < 9; i++) {
            off[i] = total;
            total += (src[i].n + 255) & ~(size_t)255;
        }
        uint8_t *d = nullptr;
        if (hipMalloc((void **)&d, total) != hipSuccess) return RAV2D_ENOMEM;
        for (int i = 0; i < 9; i++)
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
        ready[device] = true;
    }
    *out = cache[device];
    return RAV2D_OK;
}

}  // namespace r2g
