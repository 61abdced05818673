"""Zero-copy pictures: device pictures from rav2d_gpu_picture_alloc (the
PicAllocator role, picture.rs:38-41, 77-125) filtered by
rav2d_gpu_filter_planes (out of place and in place) and grained by
rav2d_gpu_apply_grain_planes (apply_grain_to_picture_mt, decode.rs:4672),
bit-exact against the oracle."""
import ctypes as C

import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O
from test_parity_deblock import _compare


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


CASES = [
    # (w, h, bitdepth, sb128, seed, tools)
    (352, 288, 10, 0, 81, R.TOOL_ALL),
    (640, 360, 8, 1, 82, R.TOOL_ALL & ~R.TOOL_GRAIN),
    (256, 136, 10, 0, 83, R.TOOL_CDEF | R.TOOL_WIENER | R.TOOL_GRAIN),  # no deblocking: LR reads the input
    (1920, 1080, 10, 1, 84, R.TOOL_ALL),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d_%db_sb%d_s%d_t%d" % (c[0], c[1], c[2], 64 << c[3],
                                                                              c[4], c[5]))
@pytest.mark.parametrize("in_place", [False, True])
def test_filter_and_grain_on_device_pictures(ctx, case, in_place):
    w, h, bd, sb128, seed, tools = case
    sf = R.SynthFrame(w, h, bd, sb128, seed, tools=tools)
    f = R.GpuFrame(ctx, sf.params)
    f.set_side_info(sf.mask_ptr, sf.n_mask, sf.lr_ptr, sf.n_lr)
    pic_in = R.Picture(ctx, w, h, bd)
    pic_out = pic_in if in_place else R.Picture(ctx, w, h, bd)
    disp = R.Picture(ctx, w, h, bd)
    try:
        fp = R.Planes()
        R._check(ctx.lib.rav2d_gpu_frame_planes(f.h, 0, C.byref(fp)), "frame_planes")
        assert list(fp.stride) == list(pic_in.planes.stride)
        pic_in.write(sf.padded_planes())  # with its picture padding
        R._check(ctx.lib.rav2d_gpu_filter_planes(ctx.h, f.h, C.byref(pic_in.planes), C.byref(pic_out.planes), None),
                 "filter_planes")
        ref = O.filter_frame(sf)
        _compare(ref, pic_out.read(), "filter_planes%s" % (" in place" if in_place else ""))
        g = sf.params.grain
        data = g if sf.params.grain_present else R.FilmGrainData()
        R._check(ctx.lib.rav2d_gpu_apply_grain_planes(ctx.h, C.byref(pic_out.planes), C.byref(disp.planes), w, h, bd,
                                                      1, C.byref(data), sf.params.grain_seed, None),
                 "apply_grain_planes")
        ctx.synchronize()
        _compare(O.apply_grain(sf, ref), disp.read(), "apply_grain_planes")
    finally:
        for p in {id(pic_in): pic_in, id(pic_out): pic_out, id(disp): disp}.values():
            p.close()
        f.close()
        sf.close()


@pytest.mark.gpu
def test_planes_with_other_strides_are_rejected(ctx):
    sf = R.SynthFrame(128, 64, 10, 0, 5, tools=R.TOOL_ALL)
    f = R.GpuFrame.from_synth(ctx, sf)
    pic = R.Picture(ctx, 256, 64, 10)  # wider picture: other strides
    try:
        with pytest.raises(R.Rav2dError, match="strides"):
            R._check(ctx.lib.rav2d_gpu_filter_planes(ctx.h, f.h, C.byref(pic.planes), C.byref(pic.planes), None),
                     "filter_planes")
    finally:
        pic.close()
        f.close()
        sf.close()
