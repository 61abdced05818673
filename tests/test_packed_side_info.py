"""Side info in the device form (Rav2dPackedFilter): a frame given packed
records filters exactly like one given Av2Filter records, and like the oracle,
through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O

CASES = [
    # (w, h, bitdepth, sb128, seed, mask_flags)
    (640, 360, 10, 0, 61, R.MASK_TXPART | R.MASK_SUBPU),
    (1920, 1080, 8, 1, 62, R.MASK_SEG),
]


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d_%db_sb%d_m%d" % (c[0], c[1], c[2], 64 << c[3], c[5]))
def test_packed_side_info_filters_like_av2filter(ctx, case):
    w, h, bd, sb128, seed, flags = case
    sf = R.SynthFrame(w, h, bd, sb128, seed, lossless_pct=5, mask_flags=flags)
    fa = R.GpuFrame.from_synth(ctx, sf)
    fb = R.GpuFrame(ctx, sf.params)
    try:
        packed = R.pack_filter(sf.mask_ptr, sf.n_mask)
        R._check(ctx.lib.rav2d_gpu_frame_set_side_info_packed(fb.h, packed.ctypes.data, sf.n_mask, sf.lr_ptr,
                                                              sf.n_lr, None), "set_side_info_packed")
        if sf.params.seg_enabled:
            fb.set_segmap(*sf.segmaps())
        fb.upload_padded(sf)
        fa.filter()
        fb.filter()
        a, b = fa.download(1), fb.download(1)
        ref = O.filter_frame(sf)
        for pl in range(3):
            assert np.array_equal(a[pl], b[pl])
            assert np.array_equal(ref[pl], b[pl])
    finally:
        fa.close()
        fb.close()
        sf.close()


@pytest.mark.gpu
def test_packed_side_info_count_checked(ctx):
    sf = R.SynthFrame(136, 72, 10, 0, seed=63)
    f = R.GpuFrame(ctx, sf.params)
    try:
        packed = R.pack_filter(sf.mask_ptr, sf.n_mask)
        rc = ctx.lib.rav2d_gpu_frame_set_side_info_packed(f.h, packed.ctypes.data, sf.n_mask + 1, None, 0, None)
        assert rc == R.RAV2D_EINVAL
    finally:
        f.close()
        sf.close()


@pytest.mark.gpu
@pytest.mark.parametrize("layout", [R.PIXEL_LAYOUT_I422, R.PIXEL_LAYOUT_I444], ids=["i422", "i444"])
def test_packed_side_info_rejected_for_full_chroma_masks(ctx, layout):
    # the packed form holds 4:2:0 chroma masks: 4:2:2 / 4:4:4 frames are refused
    # on the host, and the frame then filters from Av2Filter records as usual
    from test_layouts_422_444 import _y_stride
    from test_parity_deblock import _compare
    sf = R.SynthFrame(136, 72, 10, 0, seed=64, layout=layout)
    f = R.GpuFrame(ctx, sf.params)
    try:
        packed = R.pack_filter(sf.mask_ptr, sf.n_mask)
        rc = ctx.lib.rav2d_gpu_frame_set_side_info_packed(f.h, packed.ctypes.data, sf.n_mask, sf.lr_ptr, sf.n_lr,
                                                          None)
        assert rc == R.RAV2D_EINVAL
        assert b"4:2:0 chroma masks" in ctx.lib.rav2d_gpu_last_error()
        f.set_side_info(sf.mask_ptr, sf.n_mask, sf.lr_ptr, sf.n_lr)
        f.upload_padded(sf)
        f.filter()
        _compare(O.filter_frame(sf, y_stride=_y_stride(f)), f.download(1), "after the refused packed form")
    finally:
        f.close()
        sf.close()
