"""The oracle against the reference's own known answers and against
invariants its unit tests state (CPU only)."""
import ctypes as C

import numpy as np

import rav2d_amd as R
import oracle_lib as O


def test_dq_lookup_closed_form():
    lib = O.load()
    # quantizer.rs:31-47: 64 at 0, table value << (q-1)/24 otherwise
    assert lib.oracle_dq_lookup(0) == 64
    assert lib.oracle_dq_lookup(1) == 40
    assert lib.oracle_dq_lookup(24) == 78
    assert lib.oracle_dq_lookup(25) == 80          # 40 << 1
    tbl = [40, 41, 43, 44, 45, 47, 48, 49, 51, 52, 54, 55, 57, 59, 60, 62, 64, 66, 68, 70, 72, 74, 76, 78]
    for q in range(1, 256):
        assert lib.oracle_dq_lookup(q) == tbl[(q - 1) % 24] << ((q - 1) // 24)
    assert lib.oracle_dq_lookup(-5) == 64 and lib.oracle_dq_lookup(999) == lib.oracle_dq_lookup(255)


def test_deblock_quant_thr_kat():
    lib = O.load()
    # lf_mask.rs:830: deblock_quant_thr(0, 0) == (64 + 4) >> 9
    assert lib.oracle_deblock_quant_thr(0, 0) == (64 + 4) >> 9
    # closed form (lf_mask.rs:332-335) incl. the qmax clamp for hbd
    for hbd in (0, 1):
        for q in range(-10, 320):
            qq = min(max(q, 0), 255 + 48 * hbd)
            qq = min(max(qq, 0), 255)
            dq = 64 if qq == 0 else [40, 41, 43, 44, 45, 47, 48, 49, 51, 52, 54, 55, 57, 59, 60, 62, 64,
                                     66, 68, 70, 72, 74, 76, 78][(qq - 1) % 24] << ((qq - 1) // 24)
            assert lib.oracle_deblock_quant_thr(hbd, q) == (dq + 4) >> 9


def test_deblock_side_thr_table_endpoints():
    lib = O.load()
    # DEBLOCK_SIDE_THRESHOLDS[0] = -16 -> max(-16 + 16, 0) >> 5 = 0 at 8 bit
    assert lib.oracle_deblock_side_thr(0, 0) == 0
    # last entry 1678: (1678 + 16) >> 5 = 52 ; 10 bit: q_ind = q - 48, (1678 + (16 >> 2)) >> 3
    assert lib.oracle_deblock_side_thr(0, 295) == (1678 + 16) >> 5
    assert lib.oracle_deblock_side_thr(1, 295 + 48) == (1678 + 4) >> 3
    # monotone non-decreasing in qidx
    v = [lib.oracle_deblock_side_thr(0, q) for q in range(296)]
    assert all(a <= b for a, b in zip(v, v[1:]))


def test_uniform_frame_unchanged_by_deblock():
    """deblock.rs:2409-2427: a uniform image is left unchanged."""
    sf = R.SynthFrame(128, 128, 10, 0, seed=9, tools=R.TOOL_DEBLOCK)
    for pl in range(3):
        h, w = sf.plane_shape(pl)
        arr = np.frombuffer((C.c_uint16 * (h * w)).from_address(sf.plane_ptr(pl)), dtype=np.uint16)
        arr[:] = 512
        st = C.c_ssize_t()  # and the plane inside its picture padding
        ptr = sf._lib.r2s_plane_padded(sf._h, pl, C.byref(st))
        pad = np.frombuffer((C.c_uint8 * (st.value * (h + R.PICTURE_PAD))).from_address(ptr), dtype=np.uint16)
        pad[:] = 512
    out = O.filter_frame(sf, R.INLOOP_DEBLOCK)
    for pl in range(3):
        assert (out[pl] == 512).all()


def test_deblock_all_lossless_unchanged():
    """deblock.rs:2463-2490: lossless sides are never written."""
    sf = R.SynthFrame(256, 128, 8, 0, seed=10, tools=R.TOOL_DEBLOCK, lossless_pct=100)
    inp = sf.planes()
    out = O.filter_frame(sf, R.INLOOP_DEBLOCK)
    for a, b in zip(inp, out):
        assert np.array_equal(a, b)


def test_deblock_changes_blocky_frame():
    sf = R.SynthFrame(256, 256, 8, 0, seed=11, tools=R.TOOL_DEBLOCK)
    inp = sf.planes()
    out = O.filter_frame(sf, R.INLOOP_DEBLOCK)
    assert (inp[0] != out[0]).sum() > 1000
    # deblocking only moves samples next to transform edges: interior of 4x4 grid untouched?
    # (edges are at multiples of 4; the filter reaches at most 8 samples across)
    assert np.abs(inp[0].astype(int) - out[0]).max() < 256


def _fill(sf, values):
    """Set every sample of a plane, its picture padding included."""
    for pl in range(3):
        h, w = sf.plane_shape(pl)
        ct = C.c_uint16 if sf.bitdepth > 8 else C.c_uint8
        arr = np.frombuffer((ct * (h * w)).from_address(sf.plane_ptr(pl)), dtype=np.uint16 if sf.bitdepth > 8 else np.uint8)
        arr[:] = values[pl]
        st = C.c_ssize_t()
        ptr = sf._lib.r2s_plane_padded(sf._h, pl, C.byref(st))
        pad = np.frombuffer((C.c_uint8 * (st.value * (h + R.PICTURE_PAD))).from_address(ptr), dtype=arr.dtype)
        pad[:] = values[pl]


def test_uniform_frame_unchanged_by_cdef_and_wiener():
    """cdef.rs:1372-1392 (uniform block unchanged) and looprestoration.rs:
    2778-2820, 2958-3033, 3082-3132 (identity filters): on a flat frame every
    CDEF tap difference and every Wiener tap difference is zero, and the
    PC-Wiener filters have unit DC gain, so CDEF and NS/PC-Wiener leave it."""
    for bd in (8, 10):
        sf = R.SynthFrame(192, 128, bd, 0, seed=21 + bd, tools=R.TOOL_CDEF | R.TOOL_WIENER | R.TOOL_DEBLOCK)
        v = [300, 500, 700] if bd == 10 else [77, 120, 200]
        _fill(sf, v)
        out = O.filter_frame(sf)
        for pl in range(3):
            assert (out[pl] == v[pl]).all(), (bd, pl)


def test_lossless_frame_unchanged_by_deblock_and_restoration():
    """Lossless 4x4 blocks are never written by deblock (deblock.rs:2463-2490)
    or by LR/GDF (looprestoration.rs:2823-2862, 3036-3079, 3188-3238,
    2672-2678).  CDEF and CCSO consult the lossless masks only on
    segmented-lossless frames (cdef.rs:594-598, 894-897): the segmented case
    is test_segmentation's."""
    for bd in (8, 10):
        sf = R.SynthFrame(256, 192, bd, 0, seed=31 + bd, tools=R.TOOL_DEBLOCK | R.TOOL_WIENER | R.TOOL_GDF,
                          lossless_pct=100)
        inp = sf.planes()
        out = O.filter_frame(sf)
        for a, b in zip(inp, out):
            assert np.array_equal(a, b)


def test_zero_scaling_grain_is_a_copy():
    """filmgrain.rs:1222-1235: with a zero scaling function the noise is zero."""
    sf = R.SynthFrame(160, 96, 10, 0, seed=41, tools=R.TOOL_GRAIN)
    p = R.FrameParams.from_buffer_copy(sf.params)
    g = p.grain
    g.clip_to_restricted_range = 0
    for pl in range(3):
        for i in range(g.num_points[pl]):
            g.points[pl][i][1] = 0
    planes = sf.planes()
    disp = O.apply_grain(sf, planes, params=p)
    for a, b in zip(planes, disp):
        assert np.array_equal(a, b)


def test_grain_restricted_range_bounds():
    """filmgrain.rs:1179-1197: restricted range clips luma to [16, 235] and
    chroma to [16, 240] (scaled by 1 << (bitdepth - 8))."""
    for bd in (8, 10):
        sf = R.SynthFrame(160, 96, bd, 0, seed=43 + bd, tools=R.TOOL_GRAIN)
        p = R.FrameParams.from_buffer_copy(sf.params)
        p.grain.clip_to_restricted_range = 1
        p.grain.mc_identity = 0
        disp = O.apply_grain(sf, sf.planes(), params=p)
        s = 1 << (bd - 8)
        assert disp[0].min() >= 16 * s and disp[0].max() <= 235 * s
        for pl in (1, 2):
            assert disp[pl].min() >= 16 * s and disp[pl].max() <= 240 * s
