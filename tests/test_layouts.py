"""Monochrome (I400) frames: luma only, as the reference filters a monochrome
sequence (no chroma planes: deblock.rs:1745, 2002; cdef.rs:728; the chroma
restore bits masked off, looprestoration.rs:2401-2409; no chroma grain,
decode.rs:4693 / filmgrain.rs:745-780).  The oracle's 4:2:2 and 4:4:4 frames
are pinned here the same way: their luma, which no filter derives from chroma,
equals the 4:2:0 frame's (the GPU against the oracle in those layouts:
test_layouts_422_444.py).

Pins that do not rest on the oracle's own chroma handling: a synthetic I400
frame carries exactly the luma planes, luma side info and luma parameters of
the 4:2:0 frame of the same seed, and no luma filter reads chroma, so its
filtered and grained luma must equal the 4:2:0 frame's (checked on the oracle
on the CPU, and on the GPU at 3840x2160 without the oracle)."""
import ctypes as C

import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O
from test_parity_deblock import _compare

# (w, h, bitdepth, sb128, seed, mask_flags)
CASES = [
    (352, 288, 10, 0, 401, 0),
    (200, 136, 8, 1, 402, R.MASK_TXPART | R.MASK_SUBPU),
    (264, 136, 12, 0, 403, 0),
    (640, 360, 10, 1, 404, R.MASK_TXPART),
    (72, 40, 8, 0, 405, 0),
]


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


@pytest.mark.parametrize("case", CASES[:3], ids=lambda c: "%dx%d_%db" % c[:3])
def test_oracle_mono_luma_equals_420_luma(case):
    w, h, bd, sb, seed, mf = case
    a = R.SynthFrame(w, h, bd, sb, seed, mask_flags=mf)
    b = R.SynthFrame(w, h, bd, sb, seed, mask_flags=mf, monochrome=True)
    try:
        assert b.params.layout == R.PIXEL_LAYOUT_I400 and b.plane_shape(1) == (0, 0)
        assert np.array_equal(a.planes()[0], b.planes()[0])
        oa, ob = O.filter_frame(a), O.filter_frame(b)
        assert np.array_equal(oa[0], ob[0])
        assert (ob[0] != b.planes()[0]).mean() > 0.3  # the chain changed most samples
        assert np.array_equal(O.apply_grain(a, oa)[0], O.apply_grain(b, ob)[0])
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("layout", [R.PIXEL_LAYOUT_I422, R.PIXEL_LAYOUT_I444], ids=["i422", "i444"])
@pytest.mark.parametrize("case", CASES[:4], ids=lambda c: "%dx%d_%db" % c[:3])
def test_oracle_422_444_luma_equals_420_luma(case, layout):
    """The oracle's chroma paths for 4:2:2 / 4:4:4 (subsampling of the masks,
    CDEF blocks and the 4:2:2 direction remap, CCSO, LR stripes, grain) run
    on frames whose luma, luma side info and luma parameters are the 4:2:0
    frame's: the filtered and grained luma must be the 4:2:0 frame's, and
    every chroma plane has the layout's size and is changed by the chain."""
    w, h, bd, sb, seed, mf = case
    a = R.SynthFrame(w, h, bd, sb, seed, mask_flags=mf)
    b = R.SynthFrame(w, h, bd, sb, seed, mask_flags=mf, layout=layout)
    try:
        assert b.params.layout == layout
        ssx, ssy = int(layout != R.PIXEL_LAYOUT_I444), 0
        assert b.plane_shape(1) == (h >> ssy, w >> ssx) == R.chroma_shape(layout, w, h)
        assert np.array_equal(a.planes()[0], b.planes()[0])
        oa, ob = O.filter_frame(a), O.filter_frame(b)
        assert np.array_equal(oa[0], ob[0])
        ga, gb = O.apply_grain(a, oa), O.apply_grain(b, ob)
        assert np.array_equal(ga[0], gb[0])
        for pl in (1, 2):
            assert ob[pl].shape == b.plane_shape(pl)
            assert (ob[pl] != b.planes()[pl]).mean() > 0.3
            assert (gb[pl] != ob[pl]).mean() > 0.3
    finally:
        a.close()
        b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d_%db_sb%d" % c[:4])
def test_mono_frame_matches_oracle(ctx, case):
    w, h, bd, sb, seed, mf = case
    sf = R.SynthFrame(w, h, bd, sb, seed, mask_flags=mf, monochrome=True)
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        f.filter()
        f.apply_grain()
        ref = O.filter_frame(sf)
        _compare(ref, f.download(1), "I400 %dx%d %d-bit (output)" % (w, h, bd))
        _compare(O.apply_grain(sf, ref), f.download(2), "I400 %dx%d %d-bit (display)" % (w, h, bd))
    finally:
        f.close()
        sf.close()


@pytest.mark.gpu
def test_mono_4k_luma_equals_420_luma(ctx):
    """3840x2160 10-bit, every tool: the I400 frame's output and display luma
    equal the 4:2:0 frame's (GPU against GPU; no oracle at this size)."""
    a = R.SynthFrame(3840, 2160, 10, 1, 2024)
    b = R.SynthFrame(3840, 2160, 10, 1, 2024, monochrome=True)
    fa, fb = R.GpuFrame.from_synth(ctx, a), R.GpuFrame.from_synth(ctx, b)
    try:
        for f in (fa, fb):
            f.filter()
            f.apply_grain()
        for which in (1, 2):
            ya, yb = fa.download(which)[0], fb.download(which)[0]
            assert np.array_equal(ya, yb), "4K I400 luma differs from 4:2:0 luma (buffer %d)" % which
        assert fb.download(1)[1].shape == (0, 0)
    finally:
        fa.close()
        fb.close()
        a.close()
        b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("conf", [(0, 0), (64, 0), (32, 1)], ids=lambda c: "dbrows%d_mixed%d" % c)
def test_batch_mixes_mono_and_420(ctx, conf):
    spec = [(352, 288, 10, 0, 411, True), (640, 360, 8, 1, 412, False), (264, 136, 12, 1, 413, True),
            (320, 184, 10, 0, 414, False), (136, 72, 8, 0, 415, True)]
    sfs = [R.SynthFrame(w, h, bd, sb, s, monochrome=m) for w, h, bd, sb, s, m in spec]
    frames = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs]
    cfg = R.default_batch_config()
    cfg.deblock_rows, cfg.mixed_depths = conf
    b = R.GpuBatch(ctx, frames, cfg)
    try:
        b.run()
        b.run()
        ctx.synchronize()
        for sf, f in zip(sfs, frames):
            ref = O.filter_frame(sf)
            tag = "batch %s %d-bit" % ("I400" if sf.monochrome else "4:2:0", sf.bitdepth)
            _compare(ref, f.download(1), tag)
            _compare(O.apply_grain(sf, ref), f.download(2), tag + " display")
    finally:
        b.close()
        for f in frames:
            f.close()
        for sf in sfs:
            sf.close()


@pytest.mark.gpu
@pytest.mark.parametrize("in_place", [False, True])
def test_mono_device_pictures(ctx, in_place):
    """Zero-copy entries on I400 pictures: picture_alloc and frame_planes give no
    chroma planes, filter_planes and apply_grain_planes need none."""
    w, h, bd = 352, 288, 10
    sf = R.SynthFrame(w, h, bd, 0, 421, monochrome=True)
    f = R.GpuFrame(ctx, sf.params)
    f.set_side_info(sf.mask_ptr, sf.n_mask, sf.lr_ptr, sf.n_lr)
    pic_in = R.Picture(ctx, w, h, bd, R.PIXEL_LAYOUT_I400)
    pic_out = pic_in if in_place else R.Picture(ctx, w, h, bd, R.PIXEL_LAYOUT_I400)
    disp = R.Picture(ctx, w, h, bd, R.PIXEL_LAYOUT_I400)
    try:
        fp = R.Planes()
        R._check(ctx.lib.rav2d_gpu_frame_planes(f.h, 0, C.byref(fp)), "frame_planes")
        assert fp.data[0] and not fp.data[1] and not fp.data[2]
        assert not pic_in.planes.data[1] and not pic_in.planes.data[2]
        assert list(fp.stride) == list(pic_in.planes.stride)
        pic_in.write(sf.padded_planes())
        R._check(ctx.lib.rav2d_gpu_filter_planes(ctx.h, f.h, C.byref(pic_in.planes), C.byref(pic_out.planes), None),
                 "filter_planes")
        ref = O.filter_frame(sf)
        _compare(ref, pic_out.read(), "I400 filter_planes%s" % (" in place" if in_place else ""))
        R._check(ctx.lib.rav2d_gpu_apply_grain_planes(ctx.h, C.byref(pic_out.planes), C.byref(disp.planes), w, h, bd,
                                                      R.PIXEL_LAYOUT_I400, C.byref(sf.params.grain),
                                                      sf.params.grain_seed, None), "apply_grain_planes")
        ctx.synchronize()
        _compare(O.apply_grain(sf, ref), disp.read(), "I400 apply_grain_planes")
    finally:
        for p in {id(pic_in): pic_in, id(pic_out): pic_out, id(disp): disp}.values():
            p.close()
        f.close()
        sf.close()
