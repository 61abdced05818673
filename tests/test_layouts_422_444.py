"""4:2:2 and 4:4:4 frames on the GPU against the oracle, pass by pass (the
deblocking, CDEF with CCSO, loop restoration) and with film grain, per frame,
in batches mixed with 4:2:0 and I400 frames, and through the zero-copy entry
points.  The oracle runs with the GPU frame's luma row length: for 4:4:4 the
reference's chroma restoration reads its luma line store at column 2 x
(looprestoration.rs:1957-1959), past the frame width into the store's zero
columns and on into its next row, whose position the row length sets."""
import ctypes as C

import pytest

import rav2d_amd as R
import oracle_lib as O
from test_parity_deblock import _compare, ORACLE_DEBLOCK_COLS_ONLY

LAYOUTS = [R.PIXEL_LAYOUT_I422, R.PIXEL_LAYOUT_I444]
LNAME = {R.PIXEL_LAYOUT_I400: "i400", R.PIXEL_LAYOUT_I420: "i420", R.PIXEL_LAYOUT_I422: "i422",
         R.PIXEL_LAYOUT_I444: "i444"}

# (w, h, bitdepth, sb128, seed, mask_flags, lossless_pct)
CASES = [
    (352, 288, 10, 0, 601, 0, 0),
    (640, 360, 8, 1, 602, R.MASK_TXPART | R.MASK_SUBPU, 0),
    (264, 136, 12, 0, 603, R.MASK_SEG, 10),
    (200, 136, 10, 1, 604, R.MASK_TXPART, 5),
    (1920, 1080, 10, 1, 605, 0, 0),
]
STAGES = [("deblock", R.TOOL_DEBLOCK), ("cdef", R.TOOL_DEBLOCK | R.TOOL_CDEF | R.TOOL_CCSO),
          ("lr", R.TOOL_ALL & ~R.TOOL_GRAIN), ("all", R.TOOL_ALL)]


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


def _frame(c, layout, tools=R.TOOL_ALL):
    w, h, bd, sb, seed, mf, ll = c
    return R.SynthFrame(w, h, bd, sb, seed, tools=tools, mask_flags=mf, lossless_pct=ll, layout=layout)


def _y_stride(f):
    fp = R.Planes()
    R._check(f.lib.rav2d_gpu_frame_planes(f.h, 0, C.byref(fp)), "frame_planes")
    return fp.stride[0] // (2 if f.bitdepth > 8 else 1)


def _tag(c, layout, stage=""):
    return "%s %dx%d %d-bit sb%d m%d ll%d %s" % (LNAME[layout], c[0], c[1], c[2], 64 << c[3], c[5], c[6], stage)


@pytest.mark.gpu
@pytest.mark.parametrize("layout", LAYOUTS, ids=[LNAME[x] for x in LAYOUTS])
@pytest.mark.parametrize("c", CASES, ids=lambda c: "%dx%d_%db_sb%d_m%d" % (c[0], c[1], c[2], 64 << c[3], c[5]))
def test_passes_match_oracle(ctx, c, layout):
    for stage, tools in STAGES:
        sf = _frame(c, layout, tools)
        f = R.GpuFrame.from_synth(ctx, sf)
        try:
            ys = _y_stride(f)
            if stage == "deblock":
                f.run_pass(R.PASS_DEBLOCK_V)
                _compare(O.filter_frame(sf, R.INLOOP_DEBLOCK | ORACLE_DEBLOCK_COLS_ONLY, y_stride=ys),
                         f.download(1), _tag(c, layout, "deblock V"))
            f.filter()
            ref = O.filter_frame(sf, y_stride=ys)
            _compare(ref, f.download(1), _tag(c, layout, stage))
            if stage == "all":
                f.apply_grain()
                _compare(O.apply_grain(sf, ref), f.download(2), _tag(c, layout, "display"))
        finally:
            f.close()
            sf.close()


@pytest.mark.gpu
@pytest.mark.parametrize("conf", [(0, 0), (64, 0), (32, 1)], ids=lambda c: "dbrows%d_mixed%d" % c)
def test_batch_mixes_layouts(ctx, conf):
    spec = [(CASES[0], R.PIXEL_LAYOUT_I422), (CASES[1], R.PIXEL_LAYOUT_I444), (CASES[2], R.PIXEL_LAYOUT_I420),
            (CASES[3], R.PIXEL_LAYOUT_I400), (CASES[0], R.PIXEL_LAYOUT_I444), (CASES[2], R.PIXEL_LAYOUT_I422)]
    sfs = [_frame(c, lay) for c, lay in spec]
    frames = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs]
    cfg = R.default_batch_config()
    cfg.deblock_rows, cfg.mixed_depths = conf
    b = R.GpuBatch(ctx, frames, cfg)
    try:
        b.run()
        b.run()
        ctx.synchronize()
        for (c, lay), sf, f in zip(spec, sfs, frames):
            ref = O.filter_frame(sf, y_stride=_y_stride(f))
            _compare(ref, f.download(1), "batch " + _tag(c, lay))
            _compare(O.apply_grain(sf, ref), f.download(2), "batch " + _tag(c, lay, "display"))
    finally:
        b.close()
        for f in frames:
            f.close()
        for sf in sfs:
            sf.close()


@pytest.mark.gpu
@pytest.mark.parametrize("layout", LAYOUTS, ids=[LNAME[x] for x in LAYOUTS])
def test_device_pictures(ctx, layout):
    """filter_planes and apply_grain_planes on caller pictures of the layout."""
    c = CASES[0]
    w, h, bd = c[:3]
    sf = _frame(c, layout)
    f = R.GpuFrame(ctx, sf.params)
    f.set_side_info(sf.mask_ptr, sf.n_mask, sf.lr_ptr, sf.n_lr)
    pic_in = R.Picture(ctx, w, h, bd, layout)
    pic_out = R.Picture(ctx, w, h, bd, layout)
    disp = R.Picture(ctx, w, h, bd, layout)
    try:
        assert list(pic_in.planes.stride) == [_y_stride(f) * (2 if bd > 8 else 1), pic_in.planes.stride[1]]
        pic_in.write(sf.padded_planes())
        R._check(ctx.lib.rav2d_gpu_filter_planes(ctx.h, f.h, C.byref(pic_in.planes), C.byref(pic_out.planes), None),
                 "filter_planes")
        ref = O.filter_frame(sf, y_stride=_y_stride(f))
        _compare(ref, pic_out.read(), LNAME[layout] + " filter_planes")
        R._check(ctx.lib.rav2d_gpu_apply_grain_planes(ctx.h, C.byref(pic_out.planes), C.byref(disp.planes), w, h, bd,
                                                      layout, C.byref(sf.params.grain), sf.params.grain_seed, None),
                 "apply_grain_planes")
        ctx.synchronize()
        _compare(O.apply_grain(sf, ref), disp.read(), LNAME[layout] + " apply_grain_planes")
    finally:
        for p in (pic_in, pic_out, disp):
            p.close()
        f.close()
        sf.close()
