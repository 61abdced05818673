// Host runtime and C ABI of the gfx950 filter stage (include/rav2d_gpu.h).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "frame.h"
#include "tables.h"

namespace r2g {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

LaunchTiming &launch_timing() {
    static thread_local LaunchTiming t;
    return t;
}

static std::atomic<int> g_fail[2];

bool injected_failure(int what) {
    if (what < 0 || what > 1) return false;
    int n = g_fail[what].load();
    while (n > 0)
        if (g_fail[what].compare_exchange_weak(n, n - 1)) return true;
    return false;
}

}  // namespace r2g

using namespace r2g;

namespace {

uint32_t quant_thr(int hbd, int qidx) {
    int qmax = 255 + 48 * hbd;
    qidx = std::min(std::max(qidx, 0), qmax);
    int q = std::min(std::max(qidx, 0), 255);
    int dq = q == 0 ? 64 : ((int)R2T_DQ_LOOKUP[(q - 1) % 24] << ((q - 1) / 24));
    return (uint32_t)((dq + 4) >> 9);
}

uint32_t side_thr(int hbd, int qidx) {
    int bdm8 = 2 * hbd;
    int q_ind = std::min(std::max(qidx - 24 * bdm8, 0), 295);
    int s = R2T_DEBLOCK_SIDE_THRESHOLDS[q_ind] + ((1 << 4) >> bdm8);
    return (uint32_t)std::max(s, 0) >> (5 - bdm8);
}

// init_deblock_thr_lut_y/uv (deblock.rs:13-57), segmentation disabled.
void build_db_lut(const Rav2dFrameParams &p, int hbd, DbLut &lut) {
    for (int q = 0; q < 512; q++) {
        for (int dir = 0; dir < 2; dir++) {
            int v = q + 8 * p.db_delta_q_y[dir];
            lut.q[dir][q] = (uint16_t)quant_thr(hbd, v);
            lut.side[dir][q] = (uint16_t)side_thr(hbd, v);
        }
        int u = q + p.uac_delta + 8 * p.db_delta_q_u;
        int v = q + p.vac_delta + 8 * p.db_delta_q_v;
        lut.q[2][q] = (uint16_t)quant_thr(hbd, u);
        lut.side[2][q] = (uint16_t)side_thr(hbd, u);
        lut.q[3][q] = (uint16_t)quant_thr(hbd, v);
        lut.side[3][q] = (uint16_t)side_thr(hbd, v);
    }
}

// Bottom-of-frame crop of 32-long transform edges (deblock.rs:1572-1606),
// applied once to the library's copy for every superblock row.
void crop_bottom(std::vector<Rav2dPackedFilter> &mask, const DevGeom &g) {
    int sbh = (g.bh + (16 << g.sb128) - 1) >> (4 + g.sb128);
    for (int sby = 0; sby < sbh; sby++) {
        size_t mask_row = (size_t)((sby >> (2 - g.sb128)) * g.sb256w);
        int y64_start = sby << g.sb128;
        int y64_end = std::min((sby + 1) << g.sb128, (g.bh + 15) >> 4);
        for (int y64 = y64_start; y64 < y64_end; y64++) {
            if ((y64 + 1) * 16 + 4 <= g.bh) continue;
            int starty4 = (y64 * 16) & 0x30;
            int h4 = std::min(g.bh - y64 * 16, 16);
            int crop = starty4 + h4 - 2;
            if (crop < 0) continue;
            for (int x256 = 0; x256 < g.sb256w; x256++) {
                if (mask_row + x256 >= mask.size()) break;
                int w = std::min(64, g.bw - x256 * 64);
                uint16_t(*yv)[4] = mask[mask_row + x256].filter_y[1][crop];
                for (int i = 0; i < ((w + 15) >> 4); i++) {
                    uint16_t m = yv[3][i];
                    yv[3][i] = 0;
                    yv[2][i] |= m;
                }
            }
        }
    }
}

// init_wiener + compute_restore_planes (decode.rs:27-65) for a single frame.
void lr_frame_state(const Rav2dFrameParams &p, int &base_q, int &widx, int &ns_idx, int &restore_planes) {
    base_q = 0;
    widx = 0;
    ns_idx = 0;
    uint8_t rt = p.restoration[0].restoration_type;
    if (rt != RAV2D_RESTORATION_NONE) {
        int q = std::min(std::max(p.yac, 0), 255);
        base_q = q == 0 ? 64 : ((int)R2T_DQ_LOOKUP[(q - 1) % 24] << ((q - 1) / 24));
        widx = p.yac < 130 ? 0 : p.yac < 190 ? 1 : p.yac < 220 ? 2 : 3;
        if ((rt == RAV2D_RESTORATION_NS_WIENER || rt == RAV2D_RESTORATION_SWITCHABLE) &&
            p.restoration[0].num_classes_idx > 0)
            ns_idx = std::min(p.restoration[0].num_classes_idx - 1, 6);
    }
    int has_y = p.restoration[0].restoration_type != RAV2D_RESTORATION_NONE || p.gdf_enabled;
    int has_u = p.restoration[1].restoration_type != RAV2D_RESTORATION_NONE;
    int has_v = p.restoration[2].restoration_type != RAV2D_RESTORATION_NONE;
    restore_planes = has_y | (has_u << 1) | (has_v << 2);
    // without chroma planes only luma is restored (looprestoration.rs:2401-2409)
    if (p.layout == RAV2D_PIXEL_LAYOUT_I400) restore_planes &= 1;
}

// The stripe walk of lr_stripe (looprestoration.rs:1847-2274) for one unit
// column, emitting one work item per stripe.
// rows of the luma line store (lr_db_line: 20 rows of the plane stride, decode.rs:3818-3835;
// zero where not written, as in a fresh LoopFilterState)
constexpr int kLumaStoreRows = 20;

void lr_column_items(const Rav2dGpuFrame *f, int pl, int x, int y, int w, int row_h, const Rav2dRestorationUnit *unit,
                     uint8_t edges, bool first_sby, const int *lstore, std::vector<LrItem> &out) {
    const Rav2dFrameParams &p = f->params;
    const DevGeom &g = f->g;
    const int chroma = pl != 0, ss_ver = chroma ? g.ss_ver : 0, ss_hor = chroma ? g.ss_hor : 0;
    const int sbh = (g.bh + (16 << g.sb128) - 1) >> (4 + g.sb128);
    const int sby = (y + (y ? (8 << ss_ver) : 0)) >> (6 - ss_ver + g.sb128);
    const int sb256x = (x << ss_hor) >> 8, sb64x_idx = ((x << ss_hor) >> 6) & 3;
    int stripe_h = std::min((64 - 8 * (first_sby ? 1 : 0)) >> ss_ver, row_h - y);
    const uint32_t inloop = p.inloop_filters;
    uint8_t wtype = (inloop & RAV2D_INLOOPFILTER_WIENER) && unit ? unit->restoration_type : 0;
    const Rav2dRestorationPlane &pd = p.restoration[pl];
    int kind = LR_KIND_NONE;
    if (wtype == RAV2D_RESTORATION_NS_WIENER)
        kind = (pd.frame_filters_on && pd.num_classes != 1) ? LR_KIND_MULTI : LR_KIND_SINGLE;
    else if (wtype == RAV2D_RESTORATION_PC_WIENER)
        kind = LR_KIND_PC;
    const bool multi = kind == LR_KIND_MULTI || kind == LR_KIND_PC;
    uint16_t noskip[66] = {0};
    if (multi) {
        int r = 0;
        for (int by = y >> 2; by < (row_h >> 2) && r < 66; by++, r++) {
            size_t sbi = (size_t)g.sb256w * (by >> 6) + sb256x;
            if (sbi < f->h_mask.size()) {
                noskip[r] = f->h_mask[sbi].lr_noskip_mask[by & 63][sb64x_idx];
                if (!(edges & LRE_RIGHT) && (w & 63)) {
                    int shift = ((w >> 2) & 15) - 1;
                    uint16_t eb = (uint16_t)(noskip[r] >> shift);
                    noskip[r] |= (uint16_t)(eb << (shift + 1));
                }
            }
        }
    }
    int noskip_off = 0;
    const int hp = (g.bh * 4) >> ss_ver;
    int store_row = 0;  // first luma line-store row of the stripe (lr_stripe's llpf_off, 4 rows per stripe)
    while (y + stripe_h <= row_h) {
        bool bottom = (sby + 1) != sbh || (y + stripe_h != row_h);
        edges = (uint8_t)((edges & ~LRE_BOTTOM) | (bottom ? LRE_BOTTOM : 0));
        int inc = ((edges & LRE_TOP) && y + 8 < hp) ? 8 : 0;
        size_t sbi = (size_t)g.sb256w * ((size_t)((y << ss_ver) + inc) >> 8) + sb256x;
        LrItem it;
        memset(&it, 0, sizeof(it));
        it.plane = (int16_t)pl;
        it.x0 = (int16_t)x;
        it.w = (int16_t)w;
        it.y0 = (int16_t)y;
        it.h = (int16_t)stripe_h;
        it.edges = edges;
        it.kind = (uint8_t)kind;
        it.gdf_on = pl == 0 && (inloop & RAV2D_INLOOPFILTER_GDF) && sbi < f->h_mask.size() &&
                    f->h_mask[sbi].gdf[(((y + inc) >> 4) & 12) + sb64x_idx] != 0;
        if (kind == LR_KIND_SINGLE) {
            const int8_t *src = pd.frame_filters_on ? pd.filter[0] : unit->ns_filter[0];
            memcpy(it.filt, src, chroma ? 18 : 16);
        }
        // store rows 0, 1 above the stripe and 6, 7, 8 below it (looprestoration.rs:1957-1959,
        // 1180-1360 with ss_ver = 0)
        static const int kStoreRow[5] = {0, 1, 6, 7, 8};
        for (int k = 0; k < 5; k++) {
            const int r = store_row + kStoreRow[k];
            it.lsrc[k] = (int16_t)(r < kLumaStoreRows ? lstore[r] : -1);
        }
        if (multi)
            for (int r = 0; r < 17 && noskip_off + r < 66; r++) it.noskip[r] = noskip[noskip_off + r];
        // lossless rows (looprestoration.rs:2029-2074)
        int y4 = ((y << ss_ver) & 255) >> 2;
        int ll_rows = std::max(stripe_h >> 2, 1) + 1;
        if (sbi < f->h_mask.size()) {
            const Rav2dPackedFilter &m = f->h_mask[sbi];
            // chroma lossless rows: the packed record's 32 (4:2:0) or the full 64
            const bool full = !f->h_uvm.empty();
            const uint16_t(*uvll)[4] = full ? f->h_uvm[sbi].lossless_mask_uv : m.lossless_mask_uv;
            const int uv_rows = full ? 64 : 32;
            if (pl == 0 || ss_hor == 0) {
                const uint16_t(*src)[4] = pl == 0 ? m.lossless_mask_y : uvll;
                for (int r = 0; r < ll_rows && r < 17; r++)
                    if (y4 + r < (pl == 0 ? 64 : uv_rows)) it.ll[r] = src[y4 + r][sb64x_idx];
            } else {
                int init_y = y >> 2, end_y = (y + stripe_h) >> 2, base = y4 >> ss_ver;
                for (int yy = init_y, r = 0; yy < end_y; yy++, r++) {
                    int sy = base + (yy - init_y);
                    // rows past the chroma extent hold no chroma unit: zero
                    if (sy < uv_rows && r < ll_rows && r < 17) {
                        uint16_t lo = uvll[sy][sb64x_idx];
                        uint16_t hi = sb64x_idx + 1 < 4 ? uvll[sy][sb64x_idx + 1] : 0;
                        it.ll[r] = (uint16_t)((lo & 0xff) | ((hi & 0xff) << 8));
                    }
                }
            }
        }
        out.push_back(it);
        if (multi) noskip_off += stripe_h >> 2;
        store_row += 4;
        y += stripe_h;
        edges |= LRE_TOP;
        stripe_h = std::min(64 >> ss_ver, row_h - y);
        if (stripe_h == 0) break;
    }
}

// lr_sbrow_plane (looprestoration.rs:2276-2399): unit selection per 64 columns.
void lr_plane_items(const Rav2dGpuFrame *f, int pl, int y, int w, int h, int row_h, bool first_sby,
                    const int *lstore, std::vector<LrItem> &out) {
    const DevGeom &g = f->g;
    const int chroma = pl != 0, ss_ver = chroma ? g.ss_ver : 0, ss_hor = chroma ? g.ss_hor : 0;
    int usl = f->params.lr_unit_size[chroma ? 1 : 0];
    int unit = 1 << usl, half = unit >> 1;
    int row_y = y + ((8 >> ss_ver) * (first_sby ? 1 : 0));
    int shift_hor = 8 - ss_hor;
    int aligned = row_y & ~(unit - 1);
    if (aligned && aligned + half > h) aligned -= unit;
    aligned <<= ss_ver;
    int sb_idx = (aligned >> 8) * g.sb256w;
    int unit_base = ((aligned >> 6) & 3) << 2;
    uint8_t edges = (uint8_t)((y > 0 ? LRE_TOP : 0) | LRE_RIGHT);
    size_t lr_sb = (size_t)sb_idx, lr_u = (size_t)unit_base;
    int x = 0;
    auto unit_at = [&](size_t sbi, size_t u) -> const Rav2dRestorationUnit * {
        if (sbi < f->h_lr.size() && u < 16) return &f->h_lr[sbi].lr[pl][u];
        return nullptr;
    };
    auto emit = [&](int cx, int cw, uint8_t e) {
        const Rav2dRestorationUnit *ru = unit_at(lr_sb, lr_u);
        if (ru) {
            lr_column_items(f, pl, cx, y, cw, row_h, ru, e, first_sby, lstore, out);
        } else {
            // the reference skips the column: output equals input
            LrItem it;
            memset(&it, 0, sizeof(it));
            it.plane = (int16_t)pl;
            it.x0 = (int16_t)cx;
            it.w = (int16_t)cw;
            it.y0 = (int16_t)y;
            it.h = (int16_t)(row_h - y);
            out.push_back(it);
        }
    };
    while (x + 64 < w) {
        int next_x = x + 64;
        int nstart = next_x & ~(unit - 1);
        if (nstart && w - nstart < half) nstart -= unit;
        int next_u = unit_base + ((nstart >> (shift_hor - 2)) & 3);
        int next_sb = sb_idx + (nstart >> shift_hor);
        emit(x, 64, edges);
        lr_sb = (size_t)next_sb;
        lr_u = (size_t)next_u;
        x = next_x;
        edges |= LRE_LEFT;
    }
    edges &= ~LRE_RIGHT;
    emit(x, w - x, edges);
}

// Copy items for plane `pl`: the whole plane from the LR launch's input to
// OUT in 256x64 blocks (items of kind NONE without GDF).
void copy_items(const Rav2dGpuFrame *f, int pl, std::vector<LrItem> &out) {
    for (int y = 0; y < f->ph[pl]; y += 64)
        for (int x = 0; x < f->pw[pl]; x += 256) {
            LrItem it;
            memset(&it, 0, sizeof(it));
            it.plane = (int16_t)pl;
            it.x0 = (int16_t)x;
            it.w = (int16_t)std::min(256, f->pw[pl] - x);
            it.y0 = (int16_t)y;
            it.h = (int16_t)std::min(64, f->ph[pl] - y);
            it.kind = LR_KIND_NONE;
            out.push_back(it);
        }
}

// lr_sbrow_apply (looprestoration.rs:2401-2548) over every superblock row.
void build_lr_items(Rav2dGpuFrame *f) {
    const DevGeom &g = f->g;
    lr_frame_state(f->params, f->lr_base_q, f->lr_widx, f->lr_ns_idx, f->lr_restore_planes);
    f->lr_items.clear();
    const int rp = f->lr_restore_planes;
    const int sbh = (g.bh + (16 << g.sb128) - 1) >> (4 + g.sb128);
    // the luma line store (lr_db_line[0]) as copy_db / backup_db (deblock.rs:688-729,
    // 2192-2264) leave it for each superblock row: the deblocked luma row of each
    // store row, -1 where none was written
    int lstore[kLumaStoreRows];
    for (int &r : lstore) r = -1;
    for (int sby = 0; sby < sbh; sby++) {
        int not_last = sby + 1 < sbh;
        bool first = sby == 0;
        {
            const int H = g.bh * 4, off = 8 * (sby != 0);
            int row = (sby << (6 + g.sb128)) - off;
            const int rh = std::min((sby + 1) << (6 + g.sb128), H - 1);
            if (row < rh) {
                int sh = 64 - 8 * (row == 0), dst = 4;
                if (row > 0)
                    for (int i = 0; i < 4; i++) lstore[i] = lstore[(4 << g.sb128) + i];
                while (row + sh <= rh && dst + 4 <= kLumaStoreRows) {
                    for (int i = 0; i < 4; i++) lstore[dst + i] = row + sh - 2 + i;
                    dst += 4;
                    row += sh;
                    sh = 64;
                }
            }
        }
        for (int pl = 1; pl < 3; pl++) {
            if (!(rp & (1 << pl))) continue;
            int h = (g.bh * 4) >> g.ss_ver, w = (g.bw * 4) >> g.ss_hor;
            int next_row_y = (sby + 1) << ((6 - g.ss_ver) + g.sb128);
            int row_h = std::min(next_row_y - (8 >> g.ss_ver) * not_last, h);
            int off = (8 * (sby != 0)) >> g.ss_ver;
            int ys = (sby << ((6 - g.ss_ver) + g.sb128)) - off;
            lr_plane_items(f, pl, ys, w, h, row_h, first, lstore, f->lr_items);
        }
        if (rp & 1) {
            int h = g.bh * 4, w = g.bw * 4;
            int next_row_y = (sby + 1) << (6 + g.sb128);
            int row_h = std::min(next_row_y - 8 * not_last, h);
            int ys = (sby << (6 + g.sb128)) - 8 * (sby != 0);
            lr_plane_items(f, 0, ys, w, h, row_h, first, lstore, f->lr_items);
        }
    }
    // planes the LR launch does not restore get copy items, so that one launch
    // always leaves all three output planes written
    const Rav2dFrameParams &P = f->params;
    const bool lr_stage = rp && (P.inloop_filters & (RAV2D_INLOOPFILTER_WIENER | RAV2D_INLOOPFILTER_GDF));
    if (!lr_stage) f->lr_items.clear();
    // Items write disjoint outputs, so their order is free.  Workgroups start
    // in item order: the longest items go first and the short ones fill the
    // tail of the launch (cost = area x vector instructions per 2x2 quad of
    // each tool on gfx950, plus the staging every item pays).
    auto cost = [](const LrItem &it) {
        int c = 50 + (it.gdf_on ? 720 : 0);
        if (it.kind == LR_KIND_SINGLE)
            c += it.plane ? 350 : 140;
        else if (it.kind == LR_KIND_MULTI || it.kind == LR_KIND_PC)
            c += 300;
        return c * it.w * it.h;
    };
    std::stable_sort(f->lr_items.begin(), f->lr_items.end(),
                     [&](const LrItem &a, const LrItem &b) { return cost(a) > cost(b); });
    f->lr_n_restore = (int)f->lr_items.size();
    for (int pl = 0; pl < 3; pl++)
        if (!lr_stage || !(rp & (1 << pl))) copy_items(f, pl, f->lr_items);
}

// Launch order of the 64x64 CDEF superblocks.  The kernel gives each XCD a
// contiguous run of launch slots (xcd_item); inside a run the costliest
// superblocks go first (CDEF on: direction search and filtering; CCSO planes),
// so that the cheap ones fill the tail of the launch.
int upload_cdef_order(Rav2dGpuFrame *f) {
    const DevGeom &g = f->g;
    const Rav2dFrameParams &P = f->params;
    const int sbw = (g.bw * 4 + 63) / 64, sbh = (g.bh * 4 + 63) / 64, n = sbw * sbh;
    std::vector<int32_t> ord(n), cost(n, 0);
    const bool cdef_on = P.inloop_filters & RAV2D_INLOOPFILTER_CDEF;
    const bool ccso_on = P.inloop_filters & RAV2D_INLOOPFILTER_CCSO;
    for (int sby = 0; sby < sbh; sby++)
        for (int sbx = 0; sbx < sbw; sbx++) {
            const size_t mi = (size_t)(sby >> 2) * g.sb256w + (sbx >> 2);
            if (mi >= f->h_mask.size()) continue;
            const Rav2dPackedFilter &m = f->h_mask[mi];
            const int idx = m.cdef_idx[((sby & 3) << 2) + (sbx & 3)];
            const bool active = cdef_on && idx >= 0 && (P.cdef_y_strength[idx & 7] || P.cdef_uv_strength[idx & 7]);
            int c = active ? 8 : 0;
            for (int pl = 0; pl < 3; pl++)
                if (ccso_on && P.ccso[pl].enabled && m.ccso[pl]) c += pl ? 1 : 2;
            cost[sby * sbw + sbx] = c;
        }
    for (int i = 0; i < n; i++) ord[i] = i;
    const int q = n >> 3;
    for (int x = 0; x < 8 && q > 0; x++)
        std::stable_sort(ord.begin() + (size_t)x * q, ord.begin() + (size_t)(x + 1) * q,
                         [&](int a, int b) { return cost[a] > cost[b]; });
    HIP_TRY(hipMemcpy(f->d_cdef_order, ord.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
    return RAV2D_OK;
}

int upload_lr_items(Rav2dGpuFrame *f) {
    size_t n = f->lr_items.size();
    if (n > f->d_lr_cap) {
        if (f->d_lr_items) {
            (void)hipFree(f->d_lr_items);
            std::lock_guard<std::mutex> lk(f->ctx->mu);
            f->ctx->in_use -= f->d_lr_cap * sizeof(LrItem);
            f->bytes -= f->d_lr_cap * sizeof(LrItem);
        }
        f->d_lr_items = nullptr;
        f->d_lr_cap = 0;
        size_t cap = n + n / 4 + 16;
        HIP_TRY(hipMalloc((void **)&f->d_lr_items, cap * sizeof(LrItem)));
        std::lock_guard<std::mutex> lk(f->ctx->mu);
        f->ctx->in_use += cap * sizeof(LrItem);
        f->bytes += cap * sizeof(LrItem);
        f->d_lr_cap = cap;
    }
    if (n) HIP_TRY(hipMemcpy(f->d_lr_items, f->lr_items.data(), n * sizeof(LrItem), hipMemcpyHostToDevice));
    f->lr_dirty = false;
    return RAV2D_OK;
}

}  // namespace

namespace r2g {

// picture_has_grain (decode.rs:4641-4648)
bool data_has_grain(const Rav2dFilmGrainData &d) {
    return d.num_points[0] || d.num_points[1] || d.num_points[2] ||
           (d.clip_to_restricted_range && d.chroma_scaling_from_luma);
}

bool frame_has_grain(const Rav2dFrameParams &P) { return P.grain_present && data_has_grain(P.grain); }

}  // namespace r2g

namespace {

}  // namespace

namespace r2g {

size_t grain_scratch_bytes(int width, int height) {
    // templates, scaling LUTs (4096 entries per plane: 12-bit size), block
    // offsets at the smallest block size, and a last 256-byte line for the error word
    return 3 * 73 * 82 * 2 + 3 * 4096 + (size_t)((height + 15) / 16) * ((width + 15) / 16) * 2 + 256;
}

int grain_args_for(const Rav2dFilmGrainData &d, uint32_t seed, int bitdepth, int width, int height, uint8_t *scratch,
                   uint32_t *err, GrainArgs &a) {
    if (d.ar_coeff_lag < 0 || d.ar_coeff_lag > 3 || d.block_size < 0 || d.block_size > 1 || d.scaling_shift < 8 ||
        d.scaling_shift > 11 || d.ar_coeff_shift < 6 || d.ar_coeff_shift > 9 || d.grain_scale_shift < 0 ||
        d.grain_scale_shift > 3) {
        set_last_error("film grain parameters out of range");
        return RAV2D_EINVAL;
    }
    for (int pl = 0; pl < 3; pl++) {
        if (d.num_points[pl] < 0 || d.num_points[pl] > 14) {
            set_last_error("film grain point count out of range");
            return RAV2D_EINVAL;
        }
        // points whose x repeats or goes back are accepted: like the reference
        // (filmgrain.rs:40-46), the scaling LUT skips such segments
    }
    memset(&a, 0, sizeof(a));
    a.fg = d;
    a.seed = seed;
    a.bitdepth = bitdepth;
    a.lut = (int16_t *)scratch;
    a.scaling = scratch + 3 * 73 * 82 * 2;
    a.offsets = a.scaling + 3 * 4096;
    a.err = err;
    int bs = 16 << d.block_size;
    a.n_bands = (height + bs - 1) / bs;
    a.n_blocks = (width + bs - 1) / bs;
    a.ss_hor = a.ss_ver = 1;
    // apply_grain_row (filmgrain.rs:713, 744-747): a plane is grained only when
    // its scaling LUT exists; chroma-from-luma uses the luma LUT
    a.plane_on[0] = d.num_points[0] > 0;
    for (int uv = 1; uv < 3; uv++)
        a.plane_on[uv] = d.chroma_scaling_from_luma ? d.num_points[0] > 0 : d.num_points[uv] > 0;
    return RAV2D_OK;
}

int grain_args(Rav2dGpuFrame *f, GrainArgs &a) {
    const Rav2dFrameParams &P = f->params;
    int r = grain_args_for(P.grain, P.grain_seed, f->g.bitdepth, P.width, P.height, f->d_grain, f->d_grain_err, a);
    if (r) return r;
    for (int pl = 0; pl < 3; pl++) {
        a.src[pl] = plane_of(f, BUF_OUT, pl);
        a.dst[pl] = plane_of(f, BUF_DISP, pl);
    }
    // I400: no chroma planes to grain (decode.rs:4693, filmgrain.rs:745-780)
    if (P.layout == RAV2D_PIXEL_LAYOUT_I400) a.plane_on[1] = a.plane_on[2] = 0;
    a.ss_hor = f->g.ss_hor;
    a.ss_ver = f->g.ss_ver;
    return RAV2D_OK;
}

}  // namespace r2g

namespace {

// Grain tables depend on the frame parameters only: build them on the
// auxiliary stream so they overlap the in-loop passes.
int start_grain_setup(Rav2dGpuCtx *ctx, Rav2dGpuFrame *f, hipStream_t s) {
    GrainArgs a;
    int r = grain_args(f, a);
    if (r) return r;
    HIP_TRY(hipEventRecord(f->ev_free, s));  // earlier work on this frame comes first
    HIP_TRY(hipStreamWaitEvent(ctx->aux, f->ev_free, 0));
    launch_grain_setup(a, ctx->aux);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(f->ev_setup, ctx->aux));
    f->grain_ready = true;
    return RAV2D_OK;
}

// Waits until the frame's earlier per-frame launches (and an in-flight grain
// set-up) have finished, so that its device tables may be rewritten.
int frame_wait_idle(Rav2dGpuFrame *f) {
    if (f->done_pending) {
        HIP_TRY(hipEventSynchronize(f->ev_done));
        f->done_pending = false;
    }
    if (f->grain_ready) HIP_TRY(hipEventSynchronize(f->ev_setup));
    return RAV2D_OK;
}

int mark_done(Rav2dGpuFrame *f, hipStream_t s) {
    HIP_TRY(hipEventRecord(f->ev_done, s));
    f->done_pending = true;
    return RAV2D_OK;
}

}  // namespace

namespace r2g {

// Error flags the frame's kernels raised (the grain set-up's bounded wait):
// read after the caller's synchronization, cleared once reported.
int frame_errors(Rav2dGpuFrame *f) {
    uint32_t v = 0;
    HIP_TRY(hipMemcpy(&v, f->d_grain_err, sizeof(v), hipMemcpyDeviceToHost));
    if (v) {
        HIP_TRY(hipMemset(f->d_grain_err, 0, sizeof(v)));
        set_last_error("film grain set-up: the chroma AR filter's wait for luma timed out; display planes invalid");
        return RAV2D_EDEVICE;
    }
    return RAV2D_OK;
}

}  // namespace r2g

namespace {

int validate(const Rav2dFrameParams *p) {
    if (!p) return RAV2D_EINVAL;
    if (p->width <= 0 || p->height <= 0 || (p->width & 7) || (p->height & 7)) {
        set_last_error("frame size must be a positive multiple of 8");
        return RAV2D_EINVAL;
    }
    // the loop-restoration work items hold plane coordinates as int16_t and the
    // kernels form row offsets as 32-bit products of rows and byte strides
    if (p->width > RAV2D_MAX_FRAME_DIM || p->height > RAV2D_MAX_FRAME_DIM) {
        set_last_error("frame width and height must be at most 16384");
        return RAV2D_EINVAL;
    }
    if (p->bitdepth != 8 && p->bitdepth != 10 && p->bitdepth != 12) {
        set_last_error("bit depth must be 8, 10 or 12");
        return RAV2D_EINVAL;
    }
    if (p->layout < RAV2D_PIXEL_LAYOUT_I400 || p->layout > RAV2D_PIXEL_LAYOUT_I444) {
        set_last_error("layout must be I400, I420, I422 or I444");
        return RAV2D_EINVAL;
    }
    if (p->sb128 != 0 && p->sb128 != 1) {
        set_last_error("sb128 must be 0 or 1");
        return RAV2D_EINVAL;
    }
    return RAV2D_OK;
}

}  // namespace

namespace r2g {

DevPlane plane_of(const Rav2dGpuFrame *f, int b, int pl) {
    DevPlane d;
    d.ptr = f->ext[b][pl] ? f->ext[b][pl] : f->buf[b][pl];
    d.stride = f->stride[pl];
    d.w = f->pw[pl];
    d.h = f->ph[pl];
    return d;
}

}  // namespace r2g

namespace {

int copy_plane(Rav2dGpuFrame *f, int from, int to, int pl, hipStream_t s) {
    if (!f->pw[pl]) return RAV2D_OK;  // a plane the layout does not have
    HIP_TRY(hipMemcpy2DAsync(plane_of(f, to, pl).ptr, f->stride[pl], plane_of(f, from, pl).ptr, f->stride[pl],
                             (size_t)f->pw[pl] * f->bps, f->ph[pl], hipMemcpyDeviceToDevice, s));
    return RAV2D_OK;
}

// Plane layout of the library's pictures: RAV2D_PICTURE_PAD rows below each
// plane and, for widths that are not a multiple of 128, RAV2D_PICTURE_PAD
// samples right of it (the picture padding the deblocking reads), rows padded
// to 256 B (+256 B when the stride is a multiple of 1 KiB, keeping rows off
// one channel).
// Monochrome (I400) pictures have no chroma planes: size, stride and bytes 0.
void picture_layout(int w, int h, int bps, int layout, int64_t stride[3], size_t bytes[3], int pw[3], int ph[3]) {
    const int sx = layout_ss_hor(layout), sy = layout_ss_ver(layout);
    pw[0] = w;
    ph[0] = h;
    pw[1] = pw[2] = (w + sx) >> sx;
    ph[1] = ph[2] = (h + sy) >> sy;
    const int pad_cols = pad_columns(w);
    for (int pl = 0; pl < 3; pl++) {
        if (pl && layout == RAV2D_PIXEL_LAYOUT_I400) {
            pw[pl] = ph[pl] = 0;
            stride[pl] = 0;
            bytes[pl] = 0;
            continue;
        }
        int64_t s = ((int64_t)(pw[pl] + pad_cols) * bps + 255) & ~(int64_t)255;
        if ((s & 1023) == 0) s += 256;
        stride[pl] = s;
        bytes[pl] = (size_t)s * (ph[pl] + RAV2D_PICTURE_PAD);
    }
}

// Caller planes must carry the frame's strides (the kernels address every
// buffer of a plane with one stride) and 8-byte alignment (4-sample groups).
bool planes_match(const Rav2dGpuFrame *f, const Rav2dPlanes *p) {
    if (!p || p->stride[0] != f->stride[0] || p->stride[1] != f->stride[1]) return false;
    for (int pl = 0; pl < n_planes(f->params.layout); pl++)
        if (!p->data[pl] || ((uintptr_t)p->data[pl] & 7)) return false;
    return true;
}

}  // namespace

namespace r2g {

hipStream_t pick_stream(Rav2dGpuCtx *ctx, void *stream) { return stream ? (hipStream_t)stream : ctx->stream; }

// Which passes touch which plane.
struct Plan {
    bool dbv[3], dbh[3];
};

Plan make_plan(const Rav2dFrameParams &p) {
    Plan pl;
    bool db_on = (p.inloop_filters & RAV2D_INLOOPFILTER_DEBLOCK) && (p.db_level_y[0] || p.db_level_y[1]);
    pl.dbv[0] = db_on && p.db_level_y[0];
    pl.dbh[0] = db_on && p.db_level_y[1];
    // no chroma deblocking without chroma planes (deblock.rs:1745, 2002)
    const bool uv = p.layout != RAV2D_PIXEL_LAYOUT_I400;
    for (int c = 1; c < 3; c++) {
        bool lv = uv && (c == 1 ? p.db_level_u : p.db_level_v);
        pl.dbv[c] = db_on && lv;
        pl.dbh[c] = db_on && lv;
    }
    return pl;
}

}  // namespace r2g

namespace {

}  // namespace

namespace r2g {

// Launch arguments of the in-loop chain F0 -> OUT for one frame.  `upto`
// limits the passes (for per-pass parity): passes with index > upto are not
// planned.  Each launched pass reads a plane's current buffer and writes the
// other temporary; LR writes OUT.
int plan_chain(Rav2dGpuCtx *ctx, Rav2dGpuFrame *f, int upto, ChainPlan &cp) {
    if (!f->side_ok) {
        set_last_error("side info not set");
        return RAV2D_EINVAL;
    }
    memset(&cp, 0, sizeof(cp));
    Plan plan = make_plan(f->params);
    int cur[3] = {BUF_IN, BUF_IN, BUF_IN};
    auto next_tmp = [&](int c) { return c == BUF_T1 ? BUF_T2 : BUF_T1; };
    auto deblock = [&](DeblockArgs &a, const bool *on, bool vert) {
        a.g = f->g;
        a.mask = f->d_mask;
        a.uvm = f->d_uvm;
        a.lut = f->d_dblut;
        for (int pl = 0; pl < 3; pl++) {
            a.plane_on[pl] = on[pl];
            int out = next_tmp(cur[pl]);
            a.src[pl] = plane_of(f, cur[pl], pl);
            a.dst[pl] = plane_of(f, out, pl);
            cur[pl] = out;
        }
        a.pad_rows = RAV2D_PICTURE_PAD;
        a.pad_cols = pad_columns(f->pw[0]);
        a.map = deblock_tile_map(a, vert);
        if (f->params.seg_enabled) {
            const int bw = f->g.bw, bh = f->g.bh;
            a.seg[0] = f->d_seg;
            a.seg[1] = f->d_seg + (size_t)bw * bh;
            a.seg_stride[0] = bw;
            a.seg_stride[1] = bw >> f->g.ss_hor;
            a.qmax = 255 + 48 * f->g.hbd;
            memcpy(a.seg_dq, f->params.seg_delta_q, sizeof(a.seg_dq));
        }
    };
    if (f->params.seg_enabled && !f->seg_ok && upto >= RAV2D_PASS_DEBLOCK_V &&
        (plan.dbv[0] || plan.dbv[1] || plan.dbv[2] || plan.dbh[0] || plan.dbh[1] || plan.dbh[2])) {
        set_last_error("segmentation is enabled but no segment map was set");
        return RAV2D_EINVAL;
    }
    if (upto >= RAV2D_PASS_DEBLOCK_V && (plan.dbv[0] || plan.dbv[1] || plan.dbv[2])) {
        deblock(cp.dbv, plan.dbv, true);
        cp.on[RAV2D_PASS_DEBLOCK_V] = true;
    }
    if (upto >= RAV2D_PASS_DEBLOCK_H && (plan.dbh[0] || plan.dbh[1] || plan.dbh[2])) {
        deblock(cp.dbh, plan.dbh, false);
        cp.on[RAV2D_PASS_DEBLOCK_H] = true;
    }
    int f2buf[3] = {cur[0], cur[1], cur[2]};  // deblocked planes, read by LR at stripe edges
    // CDEF + CCSO (decode.rs:3858-3997)
    const Rav2dFrameParams &P = f->params;
    bool cdef_stage = P.seq_cdef && (P.inloop_filters & (RAV2D_INLOOPFILTER_CDEF | RAV2D_INLOOPFILTER_CCSO));
    if (upto >= RAV2D_PASS_CDEF && cdef_stage) {
        CdefArgs &a = cp.cdef;
        a.g = f->g;
        a.mask = f->d_mask;
        a.uvm = f->d_uvm;
        a.order = f->d_cdef_order;
        a.cdef_on = (P.inloop_filters & RAV2D_INLOOPFILTER_CDEF) != 0;
        a.ccso_on = (P.inloop_filters & RAV2D_INLOOPFILTER_CCSO) &&
                    (P.ccso[0].enabled || P.ccso[1].enabled || P.ccso[2].enabled);
        a.on_skiptx = P.cdef_on_skiptx != 0;
        a.damping = P.cdef_damping;
        a.any_ll = 0;
        for (int i = 0; i < 8; i++) a.any_ll |= P.seg_enabled && P.seg_lossless[i];
        a.has_chroma = P.layout != RAV2D_PIXEL_LAYOUT_I400;  // cdef.rs:728
        memcpy(a.y_strength, P.cdef_y_strength, 8);
        memcpy(a.uv_strength, P.cdef_uv_strength, 8);
        static const int8_t pos[7][2] = {{-1, 0}, {0, -1}, {-1, -1}, {-1, 1}, {-1, -2}, {1, -2}, {0, 2}};  // ccso.rs:5-13
        for (int pl = 0; pl < 3; pl++) {
            const Rav2dCcsoPlane &c = P.ccso[pl];
            CcsoDev &d = a.ccso[pl];
            d.shift = f->g.bitdepth - c.max_band_log2;
            int ext = c.ext_filter_support < 7 ? c.ext_filter_support : 0;
            d.dy = pos[ext][0];
            d.dx = pos[ext][1];
            d.quant_step = R2T_CCSO_QUANT_SZ[c.scale_idx & 3][c.quant_idx & 3];
            d.edge_clf = c.edge_clf;
            d.bo_only = c.bo_only;
            cdef_ccso_lut(d, c.filter_off, c.scale_idx & 3);
        }
        cdef_tap_offsets(a);
        for (int pl = 0; pl < 3; pl++) {
            int out = next_tmp(cur[pl]);
            a.src[pl] = plane_of(f, cur[pl], pl);
            a.dst[pl] = plane_of(f, out, pl);
            cur[pl] = out;
        }
        cp.on[RAV2D_PASS_CDEF] = true;
    }
    // loop restoration (decode.rs:3999-4057), written straight into OUT
    if (f->lr_dirty) {
        int r = upload_cdef_order(f);
        if (r) return r;
        build_lr_items(f);
        r = upload_lr_items(f);
        if (r) return r;
    }
    {
        LrArgs &a = cp.lr;
        a.g = f->g;
        a.items = f->d_lr_items;
        int r = lr_tables(ctx->device, &a.t);
        if (r) return r;
        a.cfl_ds = P.cfl_ds_filter_index;
        a.base_q = f->lr_base_q;
        a.widx = f->lr_widx;
        a.ns_idx = f->lr_ns_idx;
        a.gdf_ref = P.gdf_ref_dst_idx;
        a.gdf_qp = P.gdf_qp_idx;
        a.gdf_scale = P.gdf_scale;
        a.lstore_stride = (int32_t)(f->stride[0] / f->bps);
        for (int pl = 0; pl < 3; pl++) {
            memcpy(a.mfilt[pl], P.restoration[pl].filter, sizeof(a.mfilt[pl]));
            a.f3[pl] = plane_of(f, cur[pl], pl);
            a.f2[pl] = plane_of(f, f2buf[pl], pl);
            a.dst[pl] = plane_of(f, BUF_OUT, pl);
        }
    }
    // the LR launch restores what the stage restores and copies the rest
    if (upto >= RAV2D_PASS_LR) {
        cp.on[RAV2D_PASS_LR] = true;
        cp.n_lr = (int)f->lr_items.size();
        for (int pl = 0; pl < 3; pl++) cur[pl] = BUF_OUT;
    }
    for (int pl = 0; pl < 3; pl++) cp.final_buf[pl] = cur[pl];
    return RAV2D_OK;
}

}  // namespace r2g

namespace {

int run_chain(Rav2dGpuCtx *ctx, Rav2dGpuFrame *f, int upto, hipStream_t s) {
    ChainPlan cp;
    int r = plan_chain(ctx, f, upto, cp);
    if (r) return r;
    if (injected_failure(RAV2D_FAIL_LAUNCH)) {
        set_last_error("filter launch failed (injected)");
        return RAV2D_EDEVICE;
    }
    const int bd = f->g.bitdepth;
    if (cp.on[RAV2D_PASS_DEBLOCK_V]) {
        PassTimer pt(ctx, RAV2D_PASS_DEBLOCK_V, s);
        launch_deblock_v(cp.dbv, bd, s);
        HIP_TRY(hipGetLastError());
    }
    if (cp.on[RAV2D_PASS_DEBLOCK_H]) {
        PassTimer pt(ctx, RAV2D_PASS_DEBLOCK_H, s);
        launch_deblock_h(cp.dbh, bd, s);
        HIP_TRY(hipGetLastError());
    }
    if (cp.on[RAV2D_PASS_CDEF]) {
        PassTimer pt(ctx, RAV2D_PASS_CDEF, s);
        launch_cdef(cp.cdef, bd, s);
        HIP_TRY(hipGetLastError());
    }
    if (cp.on[RAV2D_PASS_LR]) {
        PassTimer pt(ctx, RAV2D_PASS_LR, s);
        launch_lr(cp.lr, cp.n_lr, bd, s);
        HIP_TRY(hipGetLastError());
    }
    for (int pl = 0; pl < 3; pl++) {
        if (cp.final_buf[pl] == BUF_OUT) continue;
        r = copy_plane(f, cp.final_buf[pl], BUF_OUT, pl, s);
        if (r) return r;
    }
    return RAV2D_OK;
}

}  // namespace

extern "C" {

size_t rav2d_gpu_struct_size(int which) {
    switch (which) {
        case 0: return sizeof(Rav2dAv2Filter);
        case 9: return sizeof(Rav2dPackedFilter);
        case 1: return sizeof(Rav2dAv2Restoration);
        case 2: return sizeof(Rav2dFrameParams);
        case 3: return sizeof(Rav2dFilmGrainData);
        case 4: return sizeof(Rav2dPlanes);
        case 5: return sizeof(Rav2dBatchConfig);
        case 6: return sizeof(Rav2dServeConfig);
        case 7: return sizeof(Rav2dServeResult);
        case 8: return sizeof(Rav2dWienerParams);
        default: return 0;
    }
}

const char *rav2d_gpu_version(void) { return "rav2d-gpu 0.1 (gfx950)"; }

int rav2d_gpu_debug_inject_failure(int what, int count) {
    if (what < 0 || what > 1 || count < 0) return RAV2D_EINVAL;
    g_fail[what].store(count);
    return RAV2D_OK;
}
const char *rav2d_gpu_last_error(void) { return g_last_error.c_str(); }

int rav2d_gpu_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int rav2d_gpu_ctx_create(int device, uint64_t hbm_budget_bytes, uint32_t flags, Rav2dGpuCtx **out) {
    (void)flags;
    if (!out) return RAV2D_EINVAL;
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        set_last_error("no such HIP device");
        return RAV2D_EINVAL;
    }
    HIP_TRY(hipSetDevice(device));
    Rav2dGpuCtx *c = new Rav2dGpuCtx();
    c->device = device;
    c->budget = hbm_budget_bytes;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        set_last_error("stream/event creation failed");
        delete c;
        return RAV2D_EDEVICE;
    }
    *out = c;
    return RAV2D_OK;
}

int rav2d_gpu_ctx_destroy(Rav2dGpuCtx *ctx) {
    if (!ctx) return RAV2D_EINVAL;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &r : ctx->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : ctx->free_events) (void)hipEventDestroy(e);
    for (auto &kv : ctx->allocs) (void)hipFree(kv.first);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipStreamSynchronize(ctx->aux);
    (void)hipStreamDestroy(ctx->aux);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RAV2D_OK;
}

int rav2d_gpu_ctx_synchronize(Rav2dGpuCtx *ctx) {
    if (!ctx) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RAV2D_OK;
}

uint64_t rav2d_gpu_ctx_bytes_in_use(const Rav2dGpuCtx *ctx) { return ctx ? ctx->in_use : 0; }

int rav2d_gpu_alloc(Rav2dGpuCtx *ctx, size_t bytes, void **out) {
    if (!ctx || !out || bytes == 0) return RAV2D_EINVAL;
    *out = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (ctx->budget && ctx->in_use + bytes > ctx->budget) {
            set_last_error("HBM budget exceeded");
            return RAV2D_EAGAIN;
        }
        ctx->in_use += bytes;
    }
    void *p = nullptr;
    if (injected_failure(RAV2D_FAIL_ALLOC) || hipMalloc(&p, bytes) != hipSuccess) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->in_use -= bytes;
        set_last_error("hipMalloc failed");
        return RAV2D_ENOMEM;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->allocs[p] = bytes;
    *out = p;
    return RAV2D_OK;
}

int rav2d_gpu_free(Rav2dGpuCtx *ctx, void *p) {
    if (!ctx || !p) return RAV2D_EINVAL;
    size_t bytes = 0;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->allocs.find(p);
        if (it == ctx->allocs.end()) {
            set_last_error("not a buffer of this context");
            return RAV2D_ENOENT;
        }
        bytes = it->second;
        ctx->allocs.erase(it);
    }
    HIP_TRY(hipSetDevice(ctx->device));
    // work already enqueued on the device may still use the buffer
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipFree(p));
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->in_use -= bytes;
    return RAV2D_OK;
}

int rav2d_gpu_copy(Rav2dGpuCtx *ctx, void *dst, const void *src, size_t bytes, void *stream) {
    if (!ctx || !dst || !src) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = pick_stream(ctx, stream);
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s));
    if (!stream) HIP_TRY(hipStreamSynchronize(s));
    return RAV2D_OK;
}

int rav2d_gpu_frame_create(Rav2dGpuCtx *ctx, const Rav2dFrameParams *params, Rav2dGpuFrame **out) {
    if (!ctx || !out) return RAV2D_EINVAL;
    *out = nullptr;
    int r = validate(params);
    if (r) return r;
    HIP_TRY(hipSetDevice(ctx->device));
    Rav2dGpuFrame *f = new Rav2dGpuFrame();
    f->ctx = ctx;
    f->params = *params;
    DevGeom &g = f->g;
    g.bw = ((params->width + 7) >> 3) << 1;
    g.bh = ((params->height + 7) >> 3) << 1;
    g.sb256w = (g.bw + 63) >> 6;
    g.sb256h = (g.bh + 63) >> 6;
    g.sb128 = params->sb128;
    g.bitdepth = params->bitdepth;
    g.bdmax = (1 << params->bitdepth) - 1;
    g.hbd = (params->bitdepth - 8) >> 1;  // SequenceHeader.hbd: 0, 1, 2 for 8, 10, 12 bit
    g.ss_hor = layout_ss_hor(params->layout);
    g.ss_ver = layout_ss_ver(params->layout);
    f->bps = params->bitdepth > 8 ? 2 : 1;
    size_t total = 0;
    picture_layout(params->width, params->height, f->bps, params->layout, f->stride, f->plane_bytes, f->pw, f->ph);
    for (int pl = 0; pl < 3; pl++) total += f->plane_bytes[pl];
    const size_t grain_bytes = grain_scratch_bytes(params->width, params->height);
    // side info and the LR work list are reserved here too, so that the budget
    // check covers everything the frame will hold
    const size_t n_sb = (size_t)g.sb256w * g.sb256h;
    const int sbrows = (params->height + (64 << params->sb128) - 1) / (64 << params->sb128);
    size_t lr_cap = 16;
    for (int pl = 0; pl < 3; pl++)
        lr_cap += (size_t)((f->pw[pl] + 63) / 64) * sbrows * 3 +
                  (size_t)((f->pw[pl] + 255) / 256) * ((f->ph[pl] + 63) / 64);
    const bool full_uv = params->layout == RAV2D_PIXEL_LAYOUT_I422 || params->layout == RAV2D_PIXEL_LAYOUT_I444;
    const size_t side_bytes = n_sb * (sizeof(Rav2dPackedFilter) + (full_uv ? sizeof(ChromaMasks) : 0));
    const size_t n_cdef = (size_t)((params->width + 63) / 64) * ((params->height + 63) / 64);
    size_t need = total * NBUF + sizeof(DbLut) + grain_bytes + side_bytes + lr_cap * sizeof(LrItem) +
                  n_cdef * sizeof(int32_t);
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (ctx->budget && ctx->in_use + need > ctx->budget) {
            delete f;
            set_last_error("HBM budget exceeded");
            return RAV2D_EAGAIN;
        }
        ctx->in_use += need;
    }
    f->bytes = need;
    if (injected_failure(RAV2D_FAIL_ALLOC) || hipMalloc((void **)&f->block, total * NBUF) != hipSuccess ||
        hipMalloc((void **)&f->d_dblut, sizeof(DbLut)) != hipSuccess ||
        hipMalloc((void **)&f->d_grain, grain_bytes) != hipSuccess ||
        hipMalloc((void **)&f->d_mask, n_sb * sizeof(Rav2dPackedFilter)) != hipSuccess ||
        hipMalloc((void **)&f->d_lr_items, lr_cap * sizeof(LrItem)) != hipSuccess ||
        hipMalloc((void **)&f->d_cdef_order, n_cdef * sizeof(int32_t)) != hipSuccess ||
        (full_uv && hipMalloc((void **)&f->d_uvm, n_sb * sizeof(ChromaMasks)) != hipSuccess)) {
        for (void *p : {(void *)f->block, (void *)f->d_dblut, (void *)f->d_grain, (void *)f->d_mask,
                        (void *)f->d_lr_items, (void *)f->d_cdef_order, (void *)f->d_uvm})
            if (p) (void)hipFree(p);
        {
            std::lock_guard<std::mutex> lk(ctx->mu);
            ctx->in_use -= need;
        }
        delete f;
        set_last_error("hipMalloc failed");
        return RAV2D_ENOMEM;
    }
    if (hipEventCreateWithFlags(&f->ev_setup, kStreamSyncEventFlags) != hipSuccess ||
        hipEventCreateWithFlags(&f->ev_free, kStreamSyncEventFlags) != hipSuccess ||
        hipEventCreateWithFlags(&f->ev_done, hipEventDisableTiming) != hipSuccess) {
        set_last_error("event creation failed");
        rav2d_gpu_frame_destroy(f);
        return RAV2D_EDEVICE;
    }
    // the last word of the grain scratch: error flags the set-up kernel raises
    f->d_grain_err = (uint32_t *)(f->d_grain + grain_bytes - 256);
    if (hipMemset(f->d_grain_err, 0, sizeof(uint32_t)) != hipSuccess) {
        set_last_error("hipMemset failed");
        rav2d_gpu_frame_destroy(f);
        return RAV2D_EDEVICE;
    }
    f->d_lr_cap = lr_cap;
    uint8_t *p = f->block;
    for (int b = 0; b < NBUF; b++)
        for (int pl = 0; pl < 3; pl++) {
            f->buf[b][pl] = p;
            p += f->plane_bytes[pl];
        }
    // the input planes (first in the block) start at zero, picture padding included
    if (hipMemset(f->buf[BUF_IN][0], 0, total) != hipSuccess) {
        set_last_error("hipMemset failed");
        rav2d_gpu_frame_destroy(f);
        return RAV2D_EDEVICE;
    }
    DbLut lut;
    build_db_lut(*params, g.hbd, lut);
    HIP_TRY(hipMemcpy(f->d_dblut, &lut, sizeof(lut), hipMemcpyHostToDevice));
    *out = f;
    return RAV2D_OK;
}

int rav2d_gpu_frame_destroy(Rav2dGpuFrame *f) {
    if (!f) return RAV2D_EINVAL;
    Rav2dGpuCtx *ctx = f->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->aux);
    (void)hipFree(f->block);
    (void)hipFree(f->d_dblut);
    (void)hipFree(f->d_uvm);
    if (f->ev_setup) (void)hipEventDestroy(f->ev_setup);
    if (f->ev_free) (void)hipEventDestroy(f->ev_free);
    if (f->ev_done) (void)hipEventDestroy(f->ev_done);
    if (f->d_grain) (void)hipFree(f->d_grain);
    if (f->d_mask) (void)hipFree(f->d_mask);
    if (f->d_lr_items) (void)hipFree(f->d_lr_items);
    if (f->d_cdef_order) (void)hipFree(f->d_cdef_order);
    if (f->d_seg) (void)hipFree(f->d_seg);
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->in_use -= f->bytes;
    }
    delete f;
    return RAV2D_OK;
}

int rav2d_gpu_frame_set_params(Rav2dGpuFrame *f, const Rav2dFrameParams *params) {
    if (!f) return RAV2D_EINVAL;
    int r = validate(params);
    if (r) return r;
    if (params->width != f->params.width || params->height != f->params.height ||
        params->bitdepth != f->params.bitdepth || params->layout != f->params.layout ||
        params->sb128 != f->params.sb128) {
        set_last_error("frame geometry differs");
        return RAV2D_EINVAL;
    }
    HIP_TRY(hipSetDevice(f->ctx->device));
    // earlier launches of this frame read the tables rewritten here, and a
    // grain set-up built for the old parameters must not be applied with the
    // new ones: wait for both, then drop the set-up
    r = frame_wait_idle(f);
    if (r) return r;
    f->grain_ready = false;
    f->params = *params;
    f->lr_dirty = true;
    f->gen++;
    DbLut lut;
    build_db_lut(*params, f->g.hbd, lut);
    HIP_TRY(hipMemcpy(f->d_dblut, &lut, sizeof(lut), hipMemcpyHostToDevice));
    return RAV2D_OK;
}

int rav2d_gpu_pack_filter(const Rav2dAv2Filter *in, size_t n, Rav2dPackedFilter *out) {
    if (n && (!in || !out)) return RAV2D_EINVAL;
    for (size_t i = 0; i < n; i++) {
        const Rav2dAv2Filter &a = in[i];
        Rav2dPackedFilter &b = out[i];
        memcpy(b.filter_y, a.filter_y, sizeof(b.filter_y));
        for (int d = 0; d < 2; d++)
            for (int l = 0; l < 32; l++)
                for (int k = 0; k < 5; k++)
                    for (int s = 0; s < 2; s++) b.filter_uv[d][l][k][s] = a.filter_uv[d][l][k][s];
        memcpy(b.lossless_mask_y, a.lossless_mask_y, sizeof(b.lossless_mask_y));
        memcpy(b.lossless_mask_uv, a.lossless_mask_uv, sizeof(b.lossless_mask_uv));  // rows 0..31
        memcpy(b.noskip_mask, a.noskip_mask, sizeof(b.noskip_mask));
        memcpy(b.lr_noskip_mask, a.lr_noskip_mask, sizeof(b.lr_noskip_mask));
        memcpy(b.qidx, a.qidx, sizeof(b.qidx));
        memcpy(b.cdef_idx, a.cdef_idx, sizeof(b.cdef_idx));
        memcpy(b.ccso, a.ccso, sizeof(b.ccso));
        b._pad0 = 0;
        memcpy(b.gdf, a.gdf, sizeof(b.gdf));
    }
    return RAV2D_OK;
}

namespace {

// The frame keeps `mask` (packed, bottom crop applied to this copy) on the
// host for its loop-restoration planner and CDEF order, and on the device for
// the kernels; the restoration records stay on the host.
int set_side_info_common(Rav2dGpuFrame *f, std::vector<Rav2dPackedFilter> &mask, const Rav2dAv2Restoration *lr,
                         size_t n_lr, void *stream) {
    const size_t need = (size_t)f->g.sb256w * f->g.sb256h;
    if (mask.size() != need || (lr && n_lr != need)) {
        set_last_error("side info count must equal the SB256 grid size");
        return RAV2D_EINVAL;
    }
    HIP_TRY(hipSetDevice(f->ctx->device));
    hipStream_t s = pick_stream(f->ctx, stream);
    int r = frame_wait_idle(f);
    if (r) return r;
    crop_bottom(mask, f->g);
    HIP_TRY(hipMemcpyAsync(f->d_mask, mask.data(), need * sizeof(Rav2dPackedFilter), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));  // the host copy is pageable
    f->h_mask.swap(mask);
    if (lr)
        f->h_lr.assign(lr, lr + n_lr);
    else
        f->h_lr.clear();
    f->lr_dirty = true;
    f->side_ok = true;
    f->gen++;
    return RAV2D_OK;
}

}  // namespace

int rav2d_gpu_frame_set_side_info(Rav2dGpuFrame *f, const Rav2dAv2Filter *mask, size_t n_mask,
                                  const Rav2dAv2Restoration *lr, size_t n_lr, void *stream) {
    if (!f || !mask) return RAV2D_EINVAL;
    if (n_mask != (size_t)f->g.sb256w * f->g.sb256h) {
        set_last_error("side info count must equal the SB256 grid size");
        return RAV2D_EINVAL;
    }
    std::vector<Rav2dPackedFilter> packed(n_mask);
    rav2d_gpu_pack_filter(mask, n_mask, packed.data());
    if (f->d_uvm) {
        // 4:2:2 / 4:4:4: the chroma masks go to the device at full size
        std::vector<ChromaMasks> uvm(n_mask);
        for (size_t i = 0; i < n_mask; i++) {
            memcpy(uvm[i].filter_uv, mask[i].filter_uv, sizeof(uvm[i].filter_uv));
            memcpy(uvm[i].lossless_mask_uv, mask[i].lossless_mask_uv, sizeof(uvm[i].lossless_mask_uv));
        }
        HIP_TRY(hipSetDevice(f->ctx->device));
        int r = frame_wait_idle(f);
        if (r) return r;
        HIP_TRY(hipMemcpy(f->d_uvm, uvm.data(), n_mask * sizeof(ChromaMasks), hipMemcpyHostToDevice));
        f->h_uvm.swap(uvm);
    }
    return set_side_info_common(f, packed, lr, n_lr, stream);
}

int rav2d_gpu_frame_set_side_info_packed(Rav2dGpuFrame *f, const Rav2dPackedFilter *mask, size_t n_mask,
                                         const Rav2dAv2Restoration *lr, size_t n_lr, void *stream) {
    if (!f || !mask) return RAV2D_EINVAL;
    if (n_mask != (size_t)f->g.sb256w * f->g.sb256h) {
        set_last_error("side info count must equal the SB256 grid size");
        return RAV2D_EINVAL;
    }
    if (f->d_uvm) {
        set_last_error("the packed side-info form holds 4:2:0 chroma masks: 4:2:2 and 4:4:4 frames take "
                       "rav2d_gpu_frame_set_side_info");
        return RAV2D_EINVAL;
    }
    std::vector<Rav2dPackedFilter> copy(mask, mask + n_mask);
    return set_side_info_common(f, copy, lr, n_lr, stream);
}

int rav2d_gpu_frame_set_segmap(Rav2dGpuFrame *f, const uint8_t *segmap, ptrdiff_t stride,
                               const uint8_t *segmap_uv, ptrdiff_t uv_stride, void *stream) {
    if (!f) return RAV2D_EINVAL;
    const int bw = f->g.bw, bh = f->g.bh, cw = bw >> f->g.ss_hor, ch = bh >> f->g.ss_ver;
    if (!segmap) {
        f->seg_ok = false;
        return RAV2D_OK;
    }
    if (!segmap_uv || stride < bw || uv_stride < cw) {
        set_last_error("segment maps need both planes and strides of at least the frame's 4x4 units");
        return RAV2D_EINVAL;
    }
    // packed copy, ids checked: the kernels index an 8-entry table with them
    std::vector<uint8_t> h((size_t)bw * bh + (size_t)cw * ch);
    for (int y = 0; y < bh; y++) memcpy(&h[(size_t)y * bw], segmap + (size_t)y * stride, bw);
    for (int y = 0; y < ch; y++) memcpy(&h[(size_t)bw * bh + (size_t)y * cw], segmap_uv + (size_t)y * uv_stride, cw);
    for (uint8_t id : h)
        if (id > 7) {
            set_last_error("segment ids must be 0..7");
            return RAV2D_EINVAL;
        }
    HIP_TRY(hipSetDevice(f->ctx->device));
    hipStream_t s = pick_stream(f->ctx, stream);
    int r = frame_wait_idle(f);
    if (r) return r;
    if (!f->d_seg) {
        {
            std::lock_guard<std::mutex> lk(f->ctx->mu);
            if (f->ctx->budget && f->ctx->in_use + h.size() > f->ctx->budget) {
                set_last_error("HBM budget exceeded");
                return RAV2D_EAGAIN;
            }
        }
        if (injected_failure(RAV2D_FAIL_ALLOC) || hipMalloc((void **)&f->d_seg, h.size()) != hipSuccess) {
            f->d_seg = nullptr;
            set_last_error("segment map allocation failed");
            return RAV2D_ENOMEM;
        }
        std::lock_guard<std::mutex> lk(f->ctx->mu);
        f->ctx->in_use += h.size();
        f->bytes += h.size();
    }
    HIP_TRY(hipMemcpyAsync(f->d_seg, h.data(), h.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));  // the host copy is pageable
    f->seg_ok = true;
    f->gen++;
    return RAV2D_OK;
}

int rav2d_gpu_frame_planes(const Rav2dGpuFrame *f, int which, Rav2dPlanes *out) {
    if (!f || !out || which < 0 || which > 2) return RAV2D_EINVAL;
    int b = which == 0 ? BUF_IN : which == 1 ? BUF_OUT : BUF_DISP;
    for (int pl = 0; pl < 3; pl++) out->data[pl] = f->pw[pl] ? f->buf[b][pl] : nullptr;
    out->stride[0] = f->stride[0];
    out->stride[1] = f->stride[1];
    return RAV2D_OK;
}

int rav2d_gpu_frame_upload(Rav2dGpuFrame *f, const void *y, const void *u, const void *v, void *stream) {
    const int np = f ? n_planes(f->params.layout) : 3;
    if (!f || !y || (np == 3 && (!u || !v))) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(f->ctx->device));
    hipStream_t s = pick_stream(f->ctx, stream);
    const void *src[3] = {y, u, v};
    for (int pl = 0; pl < np; pl++)
        HIP_TRY(hipMemcpy2DAsync(f->buf[BUF_IN][pl], f->stride[pl], src[pl], (size_t)f->pw[pl] * f->bps,
                                 (size_t)f->pw[pl] * f->bps, f->ph[pl], hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return RAV2D_OK;
}

int rav2d_gpu_frame_upload_padded(Rav2dGpuFrame *f, const Rav2dPlanes *src, int pad_rows, int pad_cols,
                                  void *stream) {
    if (!f || !src || pad_rows < 0 || pad_rows > RAV2D_PICTURE_PAD || pad_cols < 0 || pad_cols > RAV2D_PICTURE_PAD)
        return RAV2D_EINVAL;
    const int np = n_planes(f->params.layout);
    for (int pl = 0; pl < np; pl++)
        if (!src->data[pl] || src->stride[pl ? 1 : 0] < (ptrdiff_t)(f->pw[pl] + pad_cols) * f->bps) {
            set_last_error("upload_padded: null plane or stride below (width + pad_cols) samples");
            return RAV2D_EINVAL;
        }
    HIP_TRY(hipSetDevice(f->ctx->device));
    hipStream_t s = pick_stream(f->ctx, stream);
    const int cols = std::min(pad_cols, pad_columns(f->pw[0]));
    for (int pl = 0; pl < np; pl++)
        HIP_TRY(hipMemcpy2DAsync(f->buf[BUF_IN][pl], f->stride[pl], src->data[pl], (size_t)src->stride[pl ? 1 : 0],
                                 (size_t)(f->pw[pl] + cols) * f->bps, f->ph[pl] + pad_rows, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return RAV2D_OK;
}

int rav2d_gpu_frame_download(Rav2dGpuFrame *f, int which, void *y, void *u, void *v, void *stream) {
    const int np = f ? n_planes(f->params.layout) : 3;
    if (!f || !y || (np == 3 && (!u || !v)) || which < 0 || which > 2) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(f->ctx->device));
    hipStream_t s = pick_stream(f->ctx, stream);
    int b = which == 0 ? BUF_IN : which == 1 ? BUF_OUT : BUF_DISP;
    void *dst[3] = {y, u, v};
    for (int pl = 0; pl < np; pl++)
        HIP_TRY(hipMemcpy2DAsync(dst[pl], (size_t)f->pw[pl] * f->bps, f->buf[b][pl], f->stride[pl],
                                 (size_t)f->pw[pl] * f->bps, f->ph[pl], hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return which == 2 ? frame_errors(f) : RAV2D_OK;
}

int rav2d_gpu_filter_frame(Rav2dGpuCtx *ctx, Rav2dGpuFrame *f, void *stream) {
    if (!ctx || !f || f->ctx != ctx) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = pick_stream(ctx, stream);
    if (frame_has_grain(f->params) && !f->grain_ready) {
        int r = start_grain_setup(ctx, f, s);
        if (r) return r;
    }
    int r = run_chain(ctx, f, RAV2D_PASS_GRAIN, s);
    if (r) return r;
    return mark_done(f, s);
}

int rav2d_gpu_apply_grain(Rav2dGpuCtx *ctx, Rav2dGpuFrame *f, void *stream) {
    if (!ctx || !f || f->ctx != ctx) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = pick_stream(ctx, stream);
    if (!frame_has_grain(f->params)) {
        for (int pl = 0; pl < 3; pl++) {
            int r = copy_plane(f, BUF_OUT, BUF_DISP, pl, s);
            if (r) return r;
        }
        return mark_done(f, s);
    }
    GrainArgs a;
    int r = grain_args(f, a);
    if (r) return r;
    {
        PassTimer pt(ctx, RAV2D_PASS_GRAIN, s, false);
        if (f->grain_ready) {
            HIP_TRY(hipStreamWaitEvent(s, f->ev_setup, 0));
        } else {
            launch_grain_setup(a, s);
            HIP_TRY(hipGetLastError());
        }
        f->grain_ready = false;
        pt.arm();
        launch_grain_apply(a, f->g.bitdepth, s);
        HIP_TRY(hipGetLastError());
    }
    return mark_done(f, s);
}

int rav2d_gpu_run_pass(Rav2dGpuCtx *ctx, Rav2dGpuFrame *f, int pass, void *stream) {
    if (!ctx || !f || pass < 0 || pass > RAV2D_PASS_GRAIN) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = pick_stream(ctx, stream);
    int r = run_chain(ctx, f, pass, s);
    if (r) return r;
    return mark_done(f, s);
}

int rav2d_gpu_picture_alloc(Rav2dGpuCtx *ctx, int width, int height, int bitdepth, int layout, Rav2dPlanes *out) {
    if (!ctx || !out) return RAV2D_EINVAL;
    Rav2dFrameParams p;
    memset(&p, 0, sizeof(p));
    p.width = width;
    p.height = height;
    p.bitdepth = bitdepth;
    p.layout = layout;
    int r = validate(&p);
    if (r) return r;
    int64_t stride[3];
    size_t bytes[3];
    int pw[3], ph[3];
    picture_layout(width, height, bitdepth > 8 ? 2 : 1, layout, stride, bytes, pw, ph);
    void *mem = nullptr;
    r = rav2d_gpu_alloc(ctx, bytes[0] + bytes[1] + bytes[2], &mem);
    if (r) return r;
    // zero, picture padding included, until the caller writes it
    if (hipSetDevice(ctx->device) != hipSuccess || hipMemset(mem, 0, bytes[0] + bytes[1] + bytes[2]) != hipSuccess) {
        rav2d_gpu_free(ctx, mem);
        set_last_error("picture_alloc: hipMemset failed");
        return RAV2D_EDEVICE;
    }
    out->data[0] = mem;
    out->data[1] = bytes[1] ? (uint8_t *)mem + bytes[0] : nullptr;
    out->data[2] = bytes[2] ? (uint8_t *)mem + bytes[0] + bytes[1] : nullptr;
    out->stride[0] = stride[0];
    out->stride[1] = stride[1];
    return RAV2D_OK;
}

int rav2d_gpu_picture_free(Rav2dGpuCtx *ctx, Rav2dPlanes *p) {
    if (!ctx || !p || !p->data[0]) return RAV2D_EINVAL;
    int r = rav2d_gpu_free(ctx, p->data[0]);
    if (r) return r;
    memset(p, 0, sizeof(*p));
    return RAV2D_OK;
}

int rav2d_gpu_filter_planes(Rav2dGpuCtx *ctx, Rav2dGpuFrame *f, const Rav2dPlanes *in, const Rav2dPlanes *out,
                            void *stream) {
    if (!ctx || !f || f->ctx != ctx) return RAV2D_EINVAL;
    if (!planes_match(f, in) || !planes_match(f, out)) {
        set_last_error("filter_planes: planes need the frame's strides (rav2d_gpu_frame_planes) and 8-byte alignment");
        return RAV2D_EINVAL;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = pick_stream(ctx, stream);
    if (frame_has_grain(f->params) && !f->grain_ready) {
        int r = start_grain_setup(ctx, f, s);
        if (r) return r;
    }
    // in place (out = in): the passes that read the input after others have
    // written the output (loop restoration's stripe-edge rows when a plane is
    // not deblocked) take it from the frame's own input copy
    bool alias = false;
    const int np = n_planes(f->params.layout);
    for (int pl = 0; pl < np; pl++) alias = alias || in->data[pl] == out->data[pl];
    for (int pl = 0; pl < np; pl++) {
        if (alias)
            HIP_TRY(hipMemcpy2DAsync(f->buf[BUF_IN][pl], f->stride[pl], in->data[pl], f->stride[pl],
                                     (size_t)f->pw[pl] * f->bps, f->ph[pl], hipMemcpyDeviceToDevice, s));
        else
            f->ext[BUF_IN][pl] = (uint8_t *)in->data[pl];
        f->ext[BUF_OUT][pl] = (uint8_t *)out->data[pl];
    }
    int r = run_chain(ctx, f, RAV2D_PASS_GRAIN, s);
    memset(f->ext, 0, sizeof(f->ext));
    if (r) return r;
    return mark_done(f, s);
}

}  // extern "C"

namespace {

// The grain kernels on caller planes with the scratch at `scratch`.
int grain_planes(Rav2dGpuCtx *ctx, const Rav2dPlanes *src, const Rav2dPlanes *dst, int width, int height,
                 int bitdepth, int layout, const Rav2dFilmGrainData *data, uint32_t seed, uint8_t *scratch,
                 hipStream_t s) {
    const int bps = bitdepth > 8 ? 2 : 1;
    const int sx = layout_ss_hor(layout), sy = layout_ss_ver(layout);
    const int np = n_planes(layout), cw = np == 3 ? (width + sx) >> sx : 0, ch = np == 3 ? (height + sy) >> sy : 0;
    const int pw[3] = {width, cw, cw};
    const int ph[3] = {height, ch, ch};
    if (!data_has_grain(*data)) {
        // the display copy of a picture without grain (whatever the other fields)
        for (int pl = 0; pl < np; pl++)
            HIP_TRY(hipMemcpy2DAsync(dst->data[pl], dst->stride[pl ? 1 : 0], src->data[pl], src->stride[pl ? 1 : 0],
                                     (size_t)pw[pl] * bps, ph[pl], hipMemcpyDeviceToDevice, s));
        return RAV2D_OK;
    }
    GrainArgs a;
    int r = grain_args_for(*data, seed, bitdepth, width, height, scratch, nullptr, a);
    if (r) return r;
    for (int pl = 0; pl < 3; pl++) {
        a.src[pl] = DevPlane{(uint8_t *)src->data[pl], src->stride[pl ? 1 : 0], pw[pl], ph[pl]};
        a.dst[pl] = DevPlane{(uint8_t *)dst->data[pl], dst->stride[pl ? 1 : 0], pw[pl], ph[pl]};
    }
    if (np == 1) a.plane_on[1] = a.plane_on[2] = 0;  // I400: luma only
    a.ss_hor = sx;
    a.ss_ver = sy;
    if (!grain_init_tables()) {
        set_last_error("film grain tables: device set-up failed");
        return RAV2D_EDEVICE;
    }
    launch_grain_setup(a, s);
    HIP_TRY(hipGetLastError());
    launch_grain_apply(a, bitdepth, s);
    HIP_TRY(hipGetLastError());
    return RAV2D_OK;
}

int check_grain_planes(const Rav2dPlanes *src, const Rav2dPlanes *dst, int width, int height, int bitdepth,
                       int layout, const Rav2dFilmGrainData *data) {
    if (!src || !dst || !data) return RAV2D_EINVAL;
    Rav2dFrameParams p;
    memset(&p, 0, sizeof(p));
    p.width = width;
    p.height = height;
    p.bitdepth = bitdepth;
    p.layout = layout;
    int r = validate(&p);
    if (r) return r;
    const int bps = bitdepth > 8 ? 2 : 1;
    const int np = n_planes(layout);
    for (const Rav2dPlanes *q : {src, dst}) {
        for (int pl = 0; pl < np; pl++)
            if (!q->data[pl] || ((uintptr_t)q->data[pl] & 7)) {
                set_last_error("grain planes: plane pointers must be 8-byte aligned");
                return RAV2D_EINVAL;
            }
        const bool uv = np == 3;
        if ((q->stride[0] & 7) || (uv && (q->stride[1] & 7)) || q->stride[0] < (ptrdiff_t)width * bps ||
            (uv && q->stride[1] < (ptrdiff_t)((width + layout_ss_hor(layout)) >> layout_ss_hor(layout)) * bps) ||
            q->stride[0] >= (1 << 24) ||
            q->stride[1] >= (1 << 24)) {
            set_last_error("grain planes: strides must be multiples of 8 covering a row, below 16 MiB");
            return RAV2D_EINVAL;
        }
    }
    return RAV2D_OK;
}

}  // namespace

extern "C" {

int rav2d_gpu_apply_grain_planes(Rav2dGpuCtx *ctx, const Rav2dPlanes *src, const Rav2dPlanes *dst, int width,
                                 int height, int bitdepth, int layout, const Rav2dFilmGrainData *data, uint32_t seed,
                                 void *stream) {
    if (!ctx) return RAV2D_EINVAL;
    int r = check_grain_planes(src, dst, width, height, bitdepth, layout, data);
    if (r) return r;
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = pick_stream(ctx, stream);
    if (!data_has_grain(*data)) return grain_planes(ctx, src, dst, width, height, bitdepth, layout, data, seed, nullptr, s);
    // a scratch area of its own per call, stream-ordered
    uint8_t *scratch = nullptr;
    HIP_TRY(hipMallocAsync((void **)&scratch, grain_scratch_bytes(width, height), s));
    r = grain_planes(ctx, src, dst, width, height, bitdepth, layout, data, seed, scratch, s);
    HIP_TRY(hipFreeAsync(scratch, s));
    return r;
}

}  // extern "C"

namespace {

// Scratch of rav2d_gpu_apply_grain_planes_noalloc: one static device array per
// GPU (held until the process ends), large enough for pictures up to
// kNoAllocMaxDim x kNoAllocMaxDim.  Host code never uses the array's name as
// an address: each device's address comes from hipGetSymbolAddress, taken
// once per device with that device current.
constexpr int kNoAllocMaxDim = 8192;
constexpr size_t kNoAllocScratchBytes = 3 * 73 * 82 * 2 + 3 * 4096 +
                                        (size_t)(kNoAllocMaxDim / 16) * (kNoAllocMaxDim / 16) * 2 + 256;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[64] = {};

int noalloc_scratch(int device, uint8_t **out) {
    if (device < 0 || device >= 64) {
        set_last_error("device index out of range");
        return RAV2D_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return RAV2D_EDEVICE;
        }
        g_scratch_ptr[device] = (uint8_t *)ptr;
    }
    *out = g_scratch_ptr[device];
    return RAV2D_OK;
}

}  // namespace

extern "C" {

int rav2d_gpu_apply_grain_planes_noalloc(Rav2dGpuCtx *ctx, const Rav2dPlanes *src, const Rav2dPlanes *dst,
                                         int width, int height, int bitdepth, int layout,
                                         const Rav2dFilmGrainData *data, uint32_t seed, void *stream) {
    if (!ctx) return RAV2D_EINVAL;
    int r = check_grain_planes(src, dst, width, height, bitdepth, layout, data);
    if (r) return r;
    if (width > kNoAllocMaxDim || height > kNoAllocMaxDim ||
        grain_scratch_bytes(width, height) > kNoAllocScratchBytes) {
        set_last_error("apply_grain_planes_noalloc: pictures up to 8192 x 8192 only (static scratch)");
        return RAV2D_EINVAL;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    uint8_t *ptr = nullptr;
    r = noalloc_scratch(ctx->device, &ptr);
    if (r) return r;
    // the set-up writes every scratch entry the apply kernel reads: nothing to clear
    return grain_planes(ctx, src, dst, width, height, bitdepth, layout, data, seed, ptr, pick_stream(ctx, stream));
}

int rav2d_gpu_set_profiling(Rav2dGpuCtx *ctx, int on) {
    if (!ctx) return RAV2D_EINVAL;
    ctx->profiling = on != 0;
    return RAV2D_OK;
}

int rav2d_gpu_pass_times(Rav2dGpuCtx *ctx, double *ms, int64_t *count, int reset) {
    if (!ctx) return RAV2D_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (auto &r : ctx->recs) {
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, r.a, r.b));
        ctx->pass_ms[r.pass] += t;
        ctx->pass_count[r.pass] += 1;
        ctx->free_events.push_back(r.a);
        ctx->free_events.push_back(r.b);
    }
    ctx->recs.clear();
    for (int i = 0; i < RAV2D_NUM_PASSES; i++) {
        if (ms) ms[i] = ctx->pass_ms[i];
        if (count) count[i] = ctx->pass_count[i];
        if (reset) {
            ctx->pass_ms[i] = 0;
            ctx->pass_count[i] = 0;
        }
    }
    return RAV2D_OK;
}

int rav2d_gpu_timer_start(Rav2dGpuCtx *ctx) {
    if (!ctx) return RAV2D_EINVAL;
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    return RAV2D_OK;
}

int rav2d_gpu_timer_stop(Rav2dGpuCtx *ctx, float *ms) {
    if (!ctx || !ms) return RAV2D_EINVAL;
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    HIP_TRY(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return RAV2D_OK;
}

}  // extern "C"
