"""The synthetic workload's edge-mask producer: transform partitions (H5/V5
included) and sub-PU edges, against hand-derived masks and the reference's
mask tests (crates/rav2d/src/lf_mask.rs:659-826), plus the edge-spacing
property of every generated frame that the deblocking kernels rely on."""
import ctypes as C

import numpy as np
import pytest

import rav2d_amd as R

NONE, SPLIT, H, V, H4, V4, H5, V5 = range(8)
MASKS = C.c_uint16 * (2 * 64 * 5 * 4)


def _lib():
    lib = R.load_synth()
    lib.r2s_mask_edges_part.argtypes = [C.c_void_p] + [C.c_int] * 9 + [C.c_void_p, C.c_void_p]
    lib.r2s_mask_subpu_edges.argtypes = [C.c_void_p] + [C.c_int] * 9
    return lib


def _part(by4, bx4, w4, h4, part, tw4, th4, hlim=3, vlim=3, a=None, l=None):
    m = MASKS()
    a = a if a is not None else [0] * 64
    l = l if l is not None else [0] * 64
    ab = (C.c_uint8 * 64)(*a)
    lb = (C.c_uint8 * 64)(*l)
    _lib().r2s_mask_edges_part(C.byref(m), by4, bx4, w4, h4, part, tw4, th4, hlim, vlim, ab, lb)
    return np.frombuffer(m, dtype=np.uint16).reshape(2, 64, 5, 4).copy(), list(ab), list(lb)


def _subpu(m, by4, bx4, w4, h4, twl, thl, hsz, vsz, ds):
    buf = MASKS(*m.ravel().tolist())
    _lib().r2s_mask_subpu_edges(C.byref(buf), by4, bx4, w4, h4, twl, thl, hsz, vsz, ds)
    return np.frombuffer(buf, dtype=np.uint16).reshape(2, 64, 5, 4).copy()


def _only(m, entries):
    """Every word of m is zero except the listed (dir, line, level, word) = value."""
    exp = np.zeros_like(m)
    for k, v in entries.items():
        exp[k] = v
    np.testing.assert_array_equal(m, exp)


def test_part_none_8x8_block_has_only_outer_edges():
    # lf_mask.rs:659-680: an 8x8 transform on an 8x8 block; the contexts are 0,
    # so the outer edges take level min(1, 0) = 0 and the contexts become 1
    m, a, l = _part(0, 0, 2, 2, NONE, 2, 2)
    _only(m, {(0, 0, 0, 0): 0b11, (1, 0, 0, 0): 0b11})
    assert a[:2] == [1, 1] and l[:2] == [1, 1]     # lf_mask.rs:730-752


def test_part_none_4x4_transforms_inner_edges():
    # lf_mask.rs:682-703: 4x4 transforms on a 16x16 block: inner edges 1..3
    m, _, _ = _part(0, 0, 4, 4, NONE, 1, 1)
    exp = {(0, 0, 0, 0): 0xF, (1, 0, 0, 0): 0xF}
    for i in (1, 2, 3):
        exp[(0, i, 0, 0)] = 0xF
        exp[(1, i, 0, 0)] = 0xF
    _only(m, exp)


def test_part_h5_geometry():
    # lf_mask.rs:705-726: H5 of a 16x32 block with 8x8 transforms: quarter-height
    # rows split at half width, the middle half one transform twice as wide
    m, a, l = _part(0, 0, 4, 8, H5, 2, 2, a=[2] * 64, l=[3] * 64)
    _only(m, {
        (1, 0, 1, 0): 0xF,          # top edge, level min(1, 2)
        (0, 0, 1, 0): 0b11000011,   # left edge of the quarter rows, level 1
        (0, 0, 2, 0): 0b00111100,   # left edge of the middle half, level min(1+1, 3)
        (1, 2, 1, 0): 0xF,          # inner horizontal edges at 2 and 6
        (1, 6, 1, 0): 0xF,
        (0, 2, 1, 0): 0b11000011,   # inner vertical edge only outside the middle
    })
    assert a[:4] == [1] * 4
    assert l[:8] == [1, 1, 2, 2, 2, 2, 1, 1]


def test_part_v5_geometry():
    # lf_mask.rs:728-749 (V5), the transpose of H5
    m, a, l = _part(0, 0, 8, 4, V5, 2, 2, a=[3] * 64, l=[2] * 64)
    _only(m, {
        (0, 0, 1, 0): 0xF,
        (1, 0, 1, 0): 0b11000011,
        (1, 0, 2, 0): 0b00111100,
        (0, 2, 1, 0): 0xF,
        (0, 6, 1, 0): 0xF,
        (1, 2, 1, 0): 0b11000011,
    })
    assert l[:4] == [1] * 4
    assert a[:8] == [1, 1, 2, 2, 2, 2, 1, 1]


def test_part_h5_clamped_by_hlim_and_frame_edge():
    # the middle section's level is min(twl4c + 1, hlim); a block cut by the
    # frame edge (h4 = 5) keeps only the rows it has
    m, _, l = _part(0, 0, 4, 5, H5, 2, 2, hlim=1, vlim=3, a=[3] * 64, l=[3] * 64)
    assert m[0, 0, 1, 0] == 0b11111          # rows 0..4 all at level 1 (hlim)
    assert m[0, 0, 2, 0] == 0
    assert l[:5] == [1] * 5
    assert m[1, 6].sum() == 0                 # no edge at row 6 beyond the block


def test_subpu_hsz_only():
    # lf_mask.rs:776-788
    m = _subpu(np.zeros((2, 64, 5, 4), np.uint16), 0, 0, 8, 4, 1, 1, 4, 0, 15)
    _only(m, {(0, 4, 1, 0): 0xF, (0, 4, 4, 0): 0xF})


def test_subpu_vsz_only():
    # lf_mask.rs:790-802
    m = _subpu(np.zeros((2, 64, 5, 4), np.uint16), 0, 0, 4, 8, 1, 1, 0, 4, 15)
    _only(m, {(1, 4, 1, 0): 0xF, (1, 4, 4, 0): 0xF})


def test_subpu_both_without_ds_mask():
    # lf_mask.rs:804-824: with ds_sub_pu_mask 0 no threshold bit is set
    m = _subpu(np.zeros((2, 64, 5, 4), np.uint16), 0, 0, 8, 8, 1, 1, 4, 4, 0)
    _only(m, {(0, 4, 1, 0): 0xFF, (1, 4, 1, 0): 0xFF})


def test_subpu_marks_only_new_bits():
    # the threshold bit marks only positions not already edges at that level
    m0 = np.zeros((2, 64, 5, 4), np.uint16)
    m0[0, 4, 1, 0] = 0b0011
    m = _subpu(m0, 0, 0, 8, 4, 1, 1, 4, 0, 15)
    assert m[0, 4, 1, 0] == 0xF and m[0, 4, 4, 0] == 0b1100
    # spacing 2 inside a 16-wide block: x = 2, 4, .., 14, every one marked
    m = _subpu(np.zeros((2, 64, 5, 4), np.uint16), 0, 0, 16, 2, 0, 0, 2, 0, 15)
    assert [int(m[0, x, 0, 0]) for x in range(16)] == [0, 0] + [3, 0] * 7
    assert [int(m[0, x, 4, 0]) for x in range(16)] == [0, 0] + [3, 0] * 7
    # a 128-wide inter block (32 units): x = 16 is a multiple of 16, unmarked
    m = _subpu(np.zeros((2, 64, 5, 4), np.uint16), 0, 0, 32, 1, 2, 2, 8, 0, 15)
    assert [x for x in range(32) if m[0, x, 2, 0]] == [8, 16, 24]
    assert [x for x in range(32) if m[0, x, 4, 0]] == [8, 24]


def _edge_levels(sf, chroma, vert):
    """Frame-wide (y4, x4) -> highest edge level (-1: no edge) of one direction."""
    raw = np.frombuffer(bytes(sf.mask_bytes()), dtype=np.uint8).reshape(sf.n_mask, R.AV2FILTER_SIZE)
    off = 5120 if chroma else 0
    fm = raw[:, off:off + 5120].copy().view(np.uint16).reshape(sf.n_mask, 2, 64, 5, 4)[:, 0 if vert else 1]
    W, H = sf.width, sf.height
    bw = ((W + 7) >> 3) << 1
    sbw = (bw + 63) >> 6
    ss = 1 if chroma else 0
    pw4, ph4 = (W >> ss) // 4, (H >> ss) // 4
    n_line = 64 >> ss
    lv = -np.ones((ph4, pw4), np.int32)
    for y4 in range(ph4):
        for x4 in range(pw4):
            line, across = (x4, y4) if vert else (y4, x4)
            sb = ((y4 << ss) >> 6) * sbw + ((x4 << ss) >> 6)
            words = fm[sb, line % n_line, :, (across >> 4) & (3 >> ss)]
            bit = 1 << (across & 15)
            for lvl in (3, 2, 1, 0):
                if words[lvl] & bit:
                    lv[y4, x4] = lvl
                    break
    return lv if vert else lv.T


@pytest.mark.parametrize("flags", [R.MASK_TXPART, R.MASK_SUBPU, R.MASK_TXPART | R.MASK_SUBPU])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_generated_edges_keep_level_spacing(flags, seed):
    """An edge of level L has no other edge of its direction closer than 4 << L
    pixels on either side: what create_db_mask's level clamps give, and what
    lets the kernels filter the edges of one pass concurrently."""
    sf = R.SynthFrame(264, 136, 10, seed & 1, seed=seed, mask_flags=flags, lossless_pct=5)
    try:
        subpu_bits = 0
        for chroma in (0, 1):
            for vert in (True, False):
                lv = _edge_levels(sf, chroma, vert)   # rows: position along the edge line
                for row in lv:
                    pos = np.nonzero(row >= 0)[0]
                    for i, p in enumerate(pos):
                        reach = 1 << row[p]
                        if i > 0:
                            assert p - pos[i - 1] >= reach, (chroma, vert, p, row[p])
                        if i + 1 < len(pos):
                            assert pos[i + 1] - p >= reach, (chroma, vert, p, row[p])
        raw = np.frombuffer(bytes(sf.mask_bytes()), dtype=np.uint8).reshape(sf.n_mask, R.AV2FILTER_SIZE)
        fy = raw[:, :5120].copy().view(np.uint16).reshape(sf.n_mask, 2, 64, 5, 4)
        subpu_bits = int(np.count_nonzero(fy[:, :, :, 4]))
        if flags & R.MASK_SUBPU:
            assert subpu_bits > 0 or seed == 3  # seed 3 may draw a frame_mode 2 TIP frame
        else:
            assert subpu_bits == 0
    finally:
        sf.close()
