"""GPU deblocking vs the CPU oracle, bit-exact, through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O

ORACLE_DEBLOCK_COLS_ONLY = 1 << 8

CASES = [
    # (w, h, bitdepth, sb128, seed, lossless_pct)
    (352, 288, 8, 0, 1, 0),
    (352, 288, 10, 0, 2, 10),
    (640, 360, 8, 1, 3, 0),
    (640, 360, 10, 0, 4, 20),
    (1920, 1080, 10, 1, 5, 0),
    (1280, 720, 8, 0, 6, 5),
    (72, 40, 10, 0, 7, 0),   # smaller than one superblock, ragged edges
]

# masks beyond square intra transforms: (w, h, bitdepth, sb128, seed, lossless_pct, mask_flags)
MASK_CASES = [
    (352, 288, 8, 0, 11, 0, R.MASK_TXPART),
    (352, 288, 10, 0, 12, 5, R.MASK_SUBPU),
    (640, 360, 10, 1, 13, 0, R.MASK_TXPART | R.MASK_SUBPU),
    (264, 136, 8, 1, 15, 10, R.MASK_TXPART | R.MASK_SUBPU),
    (1920, 1080, 10, 0, 14, 3, R.MASK_TXPART | R.MASK_SUBPU),
]


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


def _compare(ref, got, tag):
    for pl in range(3):
        if not np.array_equal(ref[pl], got[pl]):
            d = np.argwhere(ref[pl] != got[pl])
            y, x = d[0]
            raise AssertionError("%s plane %d: %d mismatches, first at (y=%d, x=%d): ref %d got %d"
                                 % (tag, pl, len(d), y, x, ref[pl][y, x], got[pl][y, x]))


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d_%db_sb%d_s%d_ll%d" % (c[0], c[1], c[2], 64 << c[3], c[4], c[5]))
def test_deblock_cols_then_rows(ctx, case):
    _cols_then_rows(ctx, *case)


@pytest.mark.gpu
@pytest.mark.parametrize("case", MASK_CASES,
                         ids=lambda c: "%dx%d_%db_sb%d_s%d_ll%d_m%d" % (c[0], c[1], c[2], 64 << c[3], c[4], c[5], c[6]))
def test_deblock_partitions_and_subpu(ctx, case):
    """Transform partitions (H5/V5 included), sub-PU edges with their >> 3
    thresholds and TIP-frame masks, as create_db_mask produces them."""
    _cols_then_rows(ctx, *case)


@pytest.mark.gpu
@pytest.mark.parametrize("case", [
    (640, 360, 8, 0, 2024024),  # chroma rows below the frame
    (200, 136, 8, 0, 20),       # the same
    (328, 264, 10, 1, 38),      # chroma columns right of a width that is not a multiple of 128
], ids=lambda c: "%dx%d_%db_sb%d" % c[:4])
def test_picture_padding_is_read(ctx, case):
    """Transforms that overhang the bottom (and, for widths that are not a
    multiple of 128, the right) frame edge make the deblocking read the
    picture padding (filter_choice_bd reaches 8 samples across an edge,
    deblock.rs:81-186): the GPU reads the frame's padding as the oracle reads
    its buffer's, with the synthetic padding and with zero padding."""
    w, h, bd, sb128, seed = case
    sf = R.SynthFrame(w, h, bd, sb128, seed, tools=R.TOOL_DEBLOCK)
    f = R.GpuFrame.from_synth(ctx, sf)  # padded upload
    g = R.GpuFrame(ctx, sf.params)      # tight upload: zero padding
    try:
        g.set_side_info(sf.mask_ptr, sf.n_mask, sf.lr_ptr, sf.n_lr)
        R._check(ctx.lib.rav2d_gpu_frame_upload(g.h, sf.plane_ptr(0), sf.plane_ptr(1), sf.plane_ptr(2), None),
                 "upload")
        f.filter()
        g.filter()
        ref = O.filter_frame(sf)
        _compare(ref, f.download(1), "padded")
        # the oracle with zero padding
        lib = O.load()
        padded = [np.zeros_like(a) for a in sf.padded_planes()]
        for pl, a in enumerate(sf.planes()):
            padded[pl][:a.shape[0], :a.shape[1]] = a
        bps = padded[0].itemsize
        assert lib.oracle_filter_frame(C.byref(sf.params), sf.mask_ptr, sf.n_mask, sf.lr_ptr, sf.n_lr,
                                       padded[0].ctypes.data, padded[1].ctypes.data, padded[2].ctypes.data,
                                       padded[0].shape[1] * bps, padded[1].shape[1] * bps, R.INLOOP_ALL) == 0
        ref0 = [a[:s.shape[0], :s.shape[1]] for a, s in zip(padded, ref)]
        _compare(ref0, g.download(1), "zero padding")
        assert any(not np.array_equal(a, b) for a, b in zip(ref, ref0)), "padding never read"
    finally:
        f.close()
        g.close()
        sf.close()


def _cols_then_rows(ctx, w, h, bd, sb128, seed, ll, flags=0):
    sf = R.SynthFrame(w, h, bd, sb128, seed, tools=R.TOOL_DEBLOCK, lossless_pct=ll, mask_flags=flags)
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        # vertical edges only
        ref_v = O.filter_frame(sf, R.INLOOP_DEBLOCK | ORACLE_DEBLOCK_COLS_ONLY)
        f.run_pass(R.PASS_DEBLOCK_V)
        _compare(ref_v, f.download(1), "deblock V")
        # both passes
        ref = O.filter_frame(sf, R.INLOOP_DEBLOCK)
        f.filter()
        got = f.download(1)
        _compare(ref, got, "deblock V+H")
        if w >= 256:
            inp = sf.planes()
            assert any(not np.array_equal(a, b) for a, b in zip(inp, got)), "deblocking changed nothing"
    finally:
        f.close()
