"""The oracle's restated helpers against the reference's own unit-test known
answers (file:line of each test cited) and against what the mathematics or
the AV1 specification (whose CDEF direction search and film-grain synthesis
AV2 keeps) fixes independently of the oracle's code: brute-force direction
costs, LFSR period, interpolation through the scaling points, the grain LUT
recurrence.  CPU only."""
import ctypes as C

import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O
from oracle_lib import ptr

L = O.load()


# ---- deblock ---------------------------------------------------------------

def test_filter_choice_reference_cases():
    # deblock.rs:2386-2391: uniform 128 -> width 3
    buf = np.full(64, 128, np.uint8)
    assert L.oracle_filter_choice(ptr(buf), 0, 16, 19, 1, 3, 3, 10, 100) == 3
    # deblock.rs:2393-2406: a 50|200 step limits the width to <= 1
    buf = np.array([50] * 32 + [200] * 32, np.uint8)
    assert L.oracle_filter_choice(ptr(buf), 0, 32, 35, 1, 3, 3, 10, 20) <= 1
    # the 16-bit instantiation agrees on the same samples
    assert L.oracle_filter_choice(ptr(buf.astype(np.uint16)), 1, 32, 35, 1, 3, 3, 10, 20) == \
        L.oracle_filter_choice(ptr(buf), 0, 32, 35, 1, 3, 3, 10, 20)


def test_filter_choice_smooth_ramp_allows_full_width():
    # a linear ramp has zero second derivatives: no threshold can reject it,
    # so the widest filter the edge allows is chosen
    for mw in (1, 3, 6, 8):
        buf = np.arange(64, dtype=np.uint8) * 2
        assert L.oracle_filter_choice(ptr(buf), 0, 32, 35, 1, mw, mw, 1, 1) == mw


def test_edge_thresholds_reference_values():
    # deblock.rs:2613-2750, 2752-2853 (setup_thr_cols/rows): the edge takes the
    # rounded mean of both sides when both are nonzero, else the nonzero one
    assert L.oracle_edge_thr(20, 0) == 20
    assert L.oracle_edge_thr(20, 20) == 20
    assert L.oracle_edge_thr(30, 10) == (30 + 10 + 1) >> 1 == 20
    assert L.oracle_edge_thr(15, 5) == 10
    assert L.oracle_edge_thr(0, 30) == 30
    assert L.oracle_edge_thr(0, 0) == 0


# ---- CDEF ------------------------------------------------------------------

def test_constrain_reference_values():
    # cdef.rs:1189-1216
    assert L.oracle_cdef_constrain(0, 10, 1) == 0
    assert 0 < L.oracle_cdef_constrain(3, 10, 1) <= 3
    assert L.oracle_cdef_constrain(-5, 10, 1) < 0
    assert L.oracle_cdef_constrain(100, 10, 1) <= 10
    assert L.oracle_cdef_constrain(5, 0, 1) == 0
    # odd symmetry and the max(0, t - |d| >> s) cap, over a grid
    for d in range(-70, 71):
        for t in (0, 1, 4, 16):
            for sh in (0, 1, 3, 5):
                v = L.oracle_cdef_constrain(d, t, sh)
                assert v == -L.oracle_cdef_constrain(-d, t, sh)
                assert abs(v) <= abs(d) and abs(v) <= max(0, t - (abs(d) >> sh))


def test_adjust_strength_reference_values():
    # cdef.rs:1287-1310
    assert L.oracle_cdef_adjust_strength(100, 0) == 0
    assert L.oracle_cdef_adjust_strength(100, 32) == (100 * 4 + 8) >> 4
    assert L.oracle_cdef_adjust_strength(100, 1 << 18) > L.oracle_cdef_adjust_strength(100, 64)
    a, b, c = (L.oracle_cdef_adjust_strength(100, v) for v in (64, 256, 4096))
    assert a <= b <= c


DIV = [0, 840, 420, 280, 210, 168, 140, 120, 105]


def av1_find_dir(img, bdm8):
    """Direction search as the AV1 specification defines it: partial sums along
    the eight line families, costs normalised by line length."""
    partial = np.zeros((8, 15), np.int64)
    for i in range(8):
        for j in range(8):
            x = (int(img[i, j]) >> bdm8) - 128
            partial[0][i + j] += x
            partial[1][i + j // 2] += x
            partial[2][i] += x
            partial[3][3 + i - j // 2] += x
            partial[4][7 + i - j] += x
            partial[5][3 - i // 2 + j] += x
            partial[6][j] += x
            partial[7][i // 2 + j] += x
    cost = [0] * 8
    for i in range(8):
        cost[2] += int(partial[2][i]) ** 2
        cost[6] += int(partial[6][i]) ** 2
    cost[2] *= DIV[8]
    cost[6] *= DIV[8]
    for i in range(7):
        cost[0] += (int(partial[0][i]) ** 2 + int(partial[0][14 - i]) ** 2) * DIV[i + 1]
        cost[4] += (int(partial[4][i]) ** 2 + int(partial[4][14 - i]) ** 2) * DIV[i + 1]
    cost[0] += int(partial[0][7]) ** 2 * DIV[8]
    cost[4] += int(partial[4][7]) ** 2 * DIV[8]
    for d in (1, 3, 5, 7):
        for j in range(5):
            cost[d] += int(partial[d][3 + j]) ** 2
        cost[d] *= DIV[8]
        for j in range(3):
            cost[d] += (int(partial[d][j]) ** 2 + int(partial[d][10 - j]) ** 2) * DIV[2 * j + 2]
    best = max(range(8), key=lambda d: (cost[d], -d))
    return best, (cost[best] - cost[(best + 4) & 7]) >> 10


def _find_dir(img, hbd, bdm8):
    var = C.c_uint32()
    a = np.ascontiguousarray(img, dtype=np.uint16 if hbd else np.uint8)
    d = L.oracle_cdef_find_dir(ptr(a), hbd, 8, bdm8, C.byref(var))
    return d, var.value


def test_find_dir_reference_cases():
    # cdef.rs:1231-1266
    d, var = _find_dir(np.full((8, 8), 128), 0, 0)
    assert 0 <= d < 8 and var == 0
    img = np.where(np.arange(8)[None, :] < 4, 64, 192).repeat(8, 0)
    d, var = _find_dir(img, 0, 0)
    assert var > 0 and d == 6  # vertical stripes: columns are the lines of direction 6
    img = np.where(np.arange(8)[:, None] < 4, 64, 192).repeat(8, 1)
    d, var = _find_dir(img, 0, 0)
    assert var > 0 and d == 2  # horizontal stripes: rows


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_find_dir_equals_av1_definition(bd):
    rng = np.random.default_rng(5 + (bd - 8) // 2)
    bdm8 = bd - 8
    hbd = int(bd > 8)
    for k in range(300):
        amp = [4, 40, 120][k % 3] << bdm8
        base = rng.integers(0, (256 << bdm8) - amp)
        # oriented texture plus noise, so every direction wins sometimes
        yy, xx = np.mgrid[0:8, 0:8]
        ang = rng.uniform(0, np.pi)
        tex = np.sin((xx * np.cos(ang) + yy * np.sin(ang)) * rng.uniform(0.5, 2.0))
        img = np.clip(base + amp * (0.5 + 0.5 * tex) + rng.normal(0, amp / 10, (8, 8)), 0, (256 << bdm8) - 1)
        img = img.astype(np.int64)
        assert _find_dir(img, hbd, bdm8) == av1_find_dir(img, bdm8), k


# ---- CCSO ------------------------------------------------------------------

def test_ccso_score_reference_values():
    # ccso.rs:268-287, 400-405
    assert L.oracle_ccso_score(20, 10, 0) == 2
    assert L.oracle_ccso_score(20, 10, 1) == 1
    assert L.oracle_ccso_score(-20, 10, 0) == 0
    assert L.oracle_ccso_score(5, 10, 0) == 1 and L.oracle_ccso_score(-5, 10, 0) == 1
    assert L.oracle_ccso_score(10, 10, 0) == 1 and L.oracle_ccso_score(-10, 10, 0) == 1
    assert L.oracle_ccso_score(11, 10, 0) == 2 and L.oracle_ccso_score(-11, 10, 0) == 0


def _ccso_prep(bo_only):
    src = np.full(16 * 16, 128, np.uint8)
    left = np.full((16, 2), 128, np.uint8)
    top = np.full(32, 128, np.uint8)
    bottom = np.full(32, 128, np.uint8)
    dst = np.zeros(64, np.uint8)
    L.oracle_ccso_prep_8(ptr(dst), 8, ptr(src), 16, 0, ptr(left), ptr(top), 2, ptr(bottom), 2, 4, 0, 10, 0,
                         1 if bo_only else 0, 15, 8, 8, 0, 0)
    return dst


def test_ccso_prep_reference_values():
    # ccso.rs:289-355: band of 128 with max_band_log2 4 is 128 >> 4 = 8 (and both
    # edge classes of a flat image are 1)
    assert np.all(_ccso_prep(True) == 128 >> (8 - 4))
    full = _ccso_prep(False)
    assert np.all((full & 0x1F) == 8)
    assert np.all(full == ((1 << 5) | (1 << 3) | 8))


def test_ccso_add_reference_values():
    # ccso.rs:358-397
    dst = np.full(16, 128, np.uint8)
    idx = np.zeros(16, np.uint8)
    offs = np.full(8, 0x21, np.uint8)
    lut = np.array([0, 5, -5, 10, -10, 15, -15, 20], np.int8)
    ll = np.zeros((1, 4), np.uint16)
    L.oracle_ccso_add_8(ptr(dst), 4, ptr(idx), 4, ptr(offs), ptr(lut), 4, 4, ptr(ll))
    assert dst[0] == 133
    dst = np.full(16, 128, np.uint8)
    L.oracle_ccso_add_8(ptr(dst), 4, ptr(idx), 4, ptr(np.zeros(8, np.uint8)), ptr(np.full(8, 10, np.int8)), 4, 4,
                        ptr(np.full((1, 4), 0xFFFF, np.uint16)))
    assert np.all(dst == 128)


# ---- loop restoration ------------------------------------------------------

def test_get_qval_given_tskip_reference_values():
    # looprestoration.rs:2560-2584
    assert L.oracle_get_qval_given_tskip(0, 0, 0, 0) == 255 * -547
    assert L.oracle_get_qval_given_tskip(16, 100, 0, 0) != 0
    assert L.oracle_get_qval_given_tskip(64, 300, 1, 0) != L.oracle_get_qval_given_tskip(64, 300, 1, 2)
    # with zero inputs only the mode offset remains, for every mode
    for i, off in enumerate([-547, -21565, -573, -680]):
        assert L.oracle_get_qval_given_tskip(0, 0, i, 0) == 255 * off
    # the result is always 255 * (offset + q): a multiple of 255
    for q in (1, 17, 300, 4000):
        for t in (0, 5, 100):
            for i in range(4):
                assert L.oracle_get_qval_given_tskip(q, t, i, 2) % 255 == 0


ORIG = 6  # row buffers have their first sample at index 6


def test_backup_rows_reference_values():
    # looprestoration.rs:2595-2638
    src = np.full(32, 10, np.uint8)
    dst = np.zeros(40, np.uint8)
    L.oracle_backup_row_8(ptr(dst), ptr(src), ptr(np.zeros(32, np.uint8)), 16, 6, 0)
    assert np.all(dst[:ORIG] == 10) and np.array_equal(dst[ORIG:ORIG + 16], src[:16])
    assert np.all(dst[ORIG + 16:ORIG + 22] == src[15])
    src = np.full(32, 50, np.uint8)
    src[0], src[15] = 10, 90
    dst = np.zeros(40, np.uint8)
    L.oracle_backup_row_8(ptr(dst), ptr(src), ptr(np.full(16, 20, np.uint8)), 16, 6, 1 | 2)
    assert np.array_equal(dst[ORIG:ORIG + 16], src[:16])
    assert np.all(dst[:ORIG] == 20) and np.array_equal(dst[ORIG + 16:ORIG + 22], src[16:22])
    src = np.full(32, 42, np.uint8)
    dst = np.zeros(40, np.uint8)
    L.oracle_backup_row_lpf_8(ptr(dst), ptr(src), 16, 6, 0)
    assert np.all(dst[:ORIG] == 42) and np.array_equal(dst[ORIG:ORIG + 16], src[:16])
    assert np.all(dst[ORIG + 16:ORIG + 22] == src[15])


def _luma_rows(ds, edges=0, w=8, fill=None):
    stride = 32
    src = np.zeros(stride * 2 + 16, np.uint8)
    if fill is None:
        src[4:20] = 100
        src[4 + stride:20 + stride] = 200
    else:
        src[:] = fill
    dst = np.zeros(40, np.uint8)
    base = src[4:]  # the first sample of the row pair
    L.oracle_backup_row_luma_8(ptr(dst), ptr(base), stride, w, edges, 0 if ds else 1, 1, ds)
    return dst


def test_backup_row_luma_reference_values():
    # looprestoration.rs:2699-2775: box (2x2), vertical average, copy of the top row
    assert _luma_rows(0)[ORIG] == (100 + 200 + 100 + 200) >> 2
    assert np.all(_luma_rows(1)[ORIG:ORIG + 8] == 150)
    assert np.all(_luma_rows(2)[ORIG:ORIG + 8] == 100)
    d = _luma_rows(2, edges=1 | 2, fill=80)
    assert np.all(d[ORIG - 4:ORIG + 12] == 80)
    # without ss_ver the luma row is copied with 4 replicated samples each side
    src = np.full(64, 42, np.uint8)
    dst = np.zeros(40, np.uint8)
    L.oracle_backup_row_luma_8(ptr(dst), ptr(src), 16, 16, 0, 0, 0, 0)
    assert np.array_equal(dst[ORIG:ORIG + 16], src[:16]) and np.all(dst[ORIG - 4:ORIG] == 42)


def _gradients(rows_vals, w=4):
    rows = [np.full(32, v, np.uint8) for v in rows_vals]
    arr = (C.c_void_p * len(rows))(*[r.ctypes.data for r in rows])
    dst = np.zeros((16, 4), np.uint16)
    L.oracle_compute_gradient_row_8(ptr(dst), arr, 2, 2, w, 0)
    return dst


def test_gdf_gradients_reference_values():
    # looprestoration.rs:2640-2661
    assert np.all(_gradients([128] * 5)[:3] == 0)
    g = _gradients([100, 200, 100, 200, 100])
    assert g[0][0] > 0
    # closed form: alternating rows of 100/200 give |2b - a - c| = 200 at every
    # centre for every direction with a vertical component, 0 horizontally
    assert g[0][0] == 4 * 200 and g[0][1] == 0 and g[0][2] == 4 * 200 and g[0][3] == 4 * 200


def _gdf_add(err, ll, scale=16, bd=8):
    p = np.full(64, 128, np.uint8)
    e = np.full(64, err, np.int8)
    L.oracle_gdf_add_8(ptr(p), 8, ptr(e), 8, 8, 8, scale, ptr(np.full((2, 4), ll, np.uint16)), bd)
    return p


def test_gdf_add_reference_values():
    # looprestoration.rs:2663-2688
    assert _gdf_add(10, 0)[0] > 128
    assert np.all(_gdf_add(10, 0xFFFF) == 128)
    assert _gdf_add(-10, 0)[0] < 128
    # closed form at 8 bit: +-round(err * scale / 16)
    assert np.all(_gdf_add(10, 0) == 128 + 10) and np.all(_gdf_add(-10, 0) == 128 - 10)
    assert np.all(_gdf_add(7, 0, scale=3) == 128 + ((21 + 8) >> 4))


# ---- film grain ------------------------------------------------------------

def test_round2_reference_values():
    # filmgrain.rs:966-976
    assert L.oracle_grain_round2(10, 1) == 5
    assert L.oracle_grain_round2(11, 1) == 6
    assert L.oracle_grain_round2(7, 2) == 2
    assert L.oracle_grain_round2(-10, 1) == -5


def test_lfsr_is_maximal_length():
    # the 16-bit LFSR of the grain synthesis (taps 16, 15, 13, 4) visits all
    # 65535 non-zero states before repeating; any wrong tap shortens the cycle
    s = C.c_uint32(1)
    seen = 0
    for n in range(1, 65536):
        L.oracle_grain_random(1, C.byref(s))
        if s.value == 1:
            seen = n
            break
    assert seen == 65535
    s = C.c_uint32(0xAAAA)  # filmgrain.rs:948-954
    assert L.oracle_grain_random(8, C.byref(s)) < 256 and s.value != 0xAAAA


def _scaling(bd, pts):
    n = 256 if bd == 8 else 1 << bd
    out = np.full(n, 0xEE, np.uint8)
    a = np.array(pts, np.uint8).reshape(-1, 2) if pts else np.zeros((1, 2), np.uint8)
    L.oracle_generate_scaling(bd, ptr(a), len(pts), ptr(out))
    return out


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_generate_scaling_reference_values(bd):
    # filmgrain.rs:978-1008, 1298-1341
    sh = bd - 8
    assert np.all(_scaling(bd, []) == 0)
    assert np.all(_scaling(bd, [(128, 200)]) == 200)
    s = _scaling(bd, [(0, 0), (255, 255)])
    assert s[0] == 0 and s[-1] == 255 and 120 < s[128 << sh] < 140
    s = _scaling(bd, [(0, 0), (100, 100)])
    assert np.all(np.diff(s[:400].astype(int)) >= 0)
    if bd > 8:
        s = _scaling(bd, [(0, 10), (128, 200)])
        assert s[1] > 0 and s[2] > 0 and s[3] > 0


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_generate_scaling_interpolates_through_points(bd):
    # piecewise-linear through every point, flat outside, within one rounding
    # (8-bit) or two stacked roundings (high bit depth: the 8-bit grid, then the
    # steps between grid entries) of the exact line
    rng = np.random.default_rng(bd)
    sh = bd - 8
    for _ in range(50):
        n = int(rng.integers(2, 14))
        xs = np.sort(rng.choice(256, n, replace=False))
        ys = rng.integers(0, 256, n)
        s = _scaling(bd, list(zip(xs, ys))).astype(int)
        assert np.all(s[:xs[0] << sh] == ys[0]) and np.all(s[xs[-1] << sh:] == ys[-1])
        for i in range(n - 1):
            x0, x1 = xs[i] << sh, xs[i + 1] << sh
            assert s[x0] == ys[i]
            v = np.arange(x0, x1)
            line = ys[i] + (ys[i + 1] - ys[i]) * (v - x0) / (x1 - x0)
            assert np.all(np.abs(s[x0:x1] - line) <= (1.0 if bd == 8 else 1.5) + 1e-9)


GW, GH = 82, 73


def av1_grain_luma(seed, d, bd):
    """AV1 specification: Gaussian draws from the 16-bit LFSR, then the causal
    auto-regressive filter over rows >= 3 and columns 3..78."""
    from test_tables import T
    gauss = T["GAUSSIAN_SEQUENCE"]
    shift = 12 - bd + d.grain_scale_shift
    reg = seed
    g = np.zeros((GH, GW), np.int64)
    for y in range(GH):
        for x in range(GW):
            r = reg
            bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1
            reg = (r >> 1) | (bit << 15)
            v = gauss[(reg >> 5) & 2047]
            g[y, x] = (v + ((1 << shift) >> 1)) >> shift
    lag = d.ar_coeff_lag
    lo, hi = -(128 << (bd - 8)), (128 << (bd - 8)) - 1
    for y in range(3, GH):
        for x in range(3, GW - 3):
            pos, s = 0, 0
            for dr in range(-lag, 1):
                for dc in range(-lag, lag + 1):
                    if dr == 0 and dc == 0:
                        break
                    s += int(g[y + dr, x + dc]) * d.ar_coeffs[0][pos]
                    pos += 1
            sh = d.ar_coeff_shift
            g[y, x] = min(max(g[y, x] + ((s + ((1 << sh) >> 1)) >> sh), lo), hi)
    return g


def av1_grain_chroma(seed, uv, luma, d, bd, sub):
    from test_tables import T
    gauss = T["GAUSSIAN_SEQUENCE"]
    shift = 12 - bd + d.grain_scale_shift
    reg = seed ^ (0x49d8 if uv else 0xb524)
    w, h = (44, 38) if sub else (GW, GH)
    g = np.zeros((GH, GW), np.int64)
    for y in range(h):
        for x in range(w):
            r = reg
            bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1
            reg = (r >> 1) | (bit << 15)
            v = gauss[(reg >> 5) & 2047]
            g[y, x] = (v + ((1 << shift) >> 1)) >> shift
    lag = d.ar_coeff_lag
    lo, hi = -(128 << (bd - 8)), (128 << (bd - 8)) - 1
    c = d.ar_coeffs[1 + uv]
    for y in range(3, h):
        for x in range(3, w - 3):
            pos, s = 0, 0
            for dr in range(-lag, 1):
                for dc in range(-lag, lag + 1):
                    if dr == 0 and dc == 0:
                        if d.num_points[0] > 0:
                            ly, lx = ((y - 3) << sub) + 3, ((x - 3) << sub) + 3
                            tot = sum(int(luma[ly + i, lx + j]) for i in range(1 + sub) for j in range(1 + sub))
                            k = 2 * sub
                            s += ((tot + ((1 << k) >> 1)) >> k) * c[pos]
                        break
                    s += int(g[y + dr, x + dc]) * c[pos]
                    pos += 1
                else:
                    continue
                break
            sh = d.ar_coeff_shift
            g[y, x] = min(max(g[y, x] + ((s + ((1 << sh) >> 1)) >> sh), lo), hi)
    return g[:h, :w]


def _grain_params(rng, lag, gss):
    d = R.FilmGrainData()
    d.ar_coeff_lag = lag
    d.ar_coeff_shift = int(rng.integers(6, 10))
    d.grain_scale_shift = gss
    d.num_points[0] = 4
    for pl in range(3):
        for k in range(25 if pl else 24):
            d.ar_coeffs[pl][k] = int(rng.integers(-128, 128))
    return d


@pytest.mark.parametrize("lag,bd,gss", [(0, 8, 0), (1, 8, 1), (2, 10, 0), (3, 10, 2), (3, 8, 3), (1, 12, 0),
                                        (3, 12, 2)])
def test_grain_luts_equal_av1_definition(lag, bd, gss):
    rng = np.random.default_rng(lag * 10 + bd + gss)
    d = _grain_params(rng, lag, gss)
    seed = int(rng.integers(0, 1 << 16))
    buf_y = np.zeros((GH, GW), np.int16)
    L.oracle_generate_grain_y(ptr(buf_y), C.byref(d), seed, bd)
    ref_y = av1_grain_luma(seed, d, bd)
    assert np.array_equal(buf_y, ref_y)
    lim = 128 << (bd - 8)
    assert buf_y.min() >= -lim and buf_y.max() < lim
    for uv in (0, 1):
        buf = np.zeros((GH, GW), np.int16)
        L.oracle_generate_grain_uv(ptr(buf), ptr(buf_y), C.byref(d), seed, uv, 1, 1, bd)
        assert np.array_equal(buf[:38, :44], av1_grain_chroma(seed, uv, ref_y, d, bd, 1)), "uv %d" % uv
