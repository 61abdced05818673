// Deblocking decisions shared by the frame kernels (deblock.hip) and the
// per-family entry (dsp.hip): filter_choice_bd (crates/rav2d/src/deblock.rs:
// 81-186) on lines held in registers, and the small constant tables of
// deblock.rs:6-11.
#pragma once
#include "dev_common.h"

namespace r2g {

// Small tables indexed by per-lane values, held as packed byte constants (a
// table in memory would cost a dependent load per lookup):
// MAX_WIDTH_Y {1, 3, 6, 8}, MAX_WIDTH_UV {1, 3, 4}, Q_THRESH_MULTS
// {32, 25, 19, 19, 0, 18, 0, 17}, W_MULT {85, 51, 37, 28, 0, 20, 0, 15}.
constexpr uint32_t kMaxWidthY = 0x08060301u, kMaxWidthUV = 0x040301u;
constexpr uint64_t kQThreshMults = 0x1100120013131920ull, kWMult = 0x0f0014001c253355ull;
constexpr int kQFirst[3] = {45, 40, 32};

__device__ __forceinline__ int byte_at(uint64_t packed, int i) { return (int)((packed >> (8 * i)) & 0xff); }

// filter_choice_bd on two lines of a segment held in registers: s[k], t[k] =
// line 0 / line 3 sample at offset k - 8 across the edge (positive side k >= 8).
__device__ __forceinline__ int filter_choice(const int (&s)[16], const int (&t)[16], int mwn, int mwp, uint32_t q_thr,
                                             uint32_t side_thr) {
    uint32_t sd[4];
#pragma unroll
    for (int d = -2; d < 2; d++) {
        uint32_t ds = (uint32_t)abs(s[d + 7] - 2 * s[d + 8] + s[d + 9]);
        uint32_t dt = (uint32_t)abs(t[d + 7] - 2 * t[d + 8] + t[d + 9]);
        sd[d + 2] = (ds + dt + 1) >> 1;
    }
    uint32_t high = dmax(sd[0], sd[3]);
    if (high > side_thr) return 0;
    if (mwp == 1) return 1;
    uint32_t transition = sd[1] + sd[2];
    if (high > (side_thr >> 2)) return 1;
    if (transition > q_thr * 4) return 1;
    if (high > (side_thr >> 3)) return 2;
    if (transition > q_thr * 3) return 2;
    uint32_t end_thr = (side_thr * 3) >> 4;
    if (mwn >= 3) {
        uint32_t ds = (uint32_t)abs(s[7] - s[4] - 3 * (s[7] - s[6]));
        uint32_t dt = (uint32_t)abs(t[7] - t[4] - 3 * (t[7] - t[6]));
        if (((ds + dt + 1) >> 1) > end_thr) return 2;
    }
    {
        uint32_t ds = (uint32_t)abs(s[8] - s[11] - 3 * (s[8] - s[9]));
        uint32_t dt = (uint32_t)abs(t[8] - t[11] - 3 * (t[8] - t[9]));
        if (((ds + dt + 1) >> 1) > end_thr) return 2;
    }
    if (mwp == 3) return 3;
    transition <<= 4;
    int prev = 3;
#pragma unroll
    for (int dist = 4; dist <= 8; dist += 2) {  // widths above 3 are at most 8
        if (dist > mwp) break;
        uint32_t q4 = q_thr * (uint32_t)kQFirst[(dist - 4) >> 1];
        uint32_t e4 = (side_thr * (uint32_t)dist) >> 4;
        if (transition > q4) return prev;
        const int d2 = dist < 7 ? dist : 7;
        if (mwn >= d2) {
            uint32_t ds = (uint32_t)abs(s[7] - s[7 - d2] - d2 * (s[7] - s[6]));
            uint32_t dt = (uint32_t)abs(t[7] - t[7 - d2] - d2 * (t[7] - t[6]));
            if (((ds + dt + 1) >> 1) > e4) return prev;
        }
        uint32_t ds = (uint32_t)abs(s[8] - s[8 + d2] - d2 * (s[8] - s[9]));
        uint32_t dt = (uint32_t)abs(t[8] - t[8 + d2] - d2 * (t[8] - t[9]));
        if (((ds + dt + 1) >> 1) > e4) return prev;
        prev = dist;
    }
    return mwp;
}

}  // namespace r2g
