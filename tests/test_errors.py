"""Argument checks of the C ABI: invalid frame parameters and film-grain data
are rejected on the host with RAV2D_EINVAL and a message, before anything is
launched, and a rejected call leaves the frame usable.

The kernels rely on these checks: planes are multiples of 4 samples wide and
tall (frames are multiples of 8) and at most 16384 (loop-restoration work
items hold coordinates as int16_t), samples are 8- or 10-bit, and each plane
has at most 14 scaling points.  Mirrors the parameter validation of the
reference's frame header and film-grain parsing (rav2d's Rav2dFilmGrainData
carries at most 14 points per plane, filmgrain.rs:22-113).  Scaling points
whose x repeats or goes back are not an error: the reference skips those
segments (filmgrain.rs:40-46) and so does the library."""
import ctypes as C

import pytest

import rav2d_amd as R


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


def _params(sf):
    return R.FrameParams.from_buffer_copy(sf.params)


@pytest.mark.gpu
@pytest.mark.parametrize("field,value,msg", [
    ("width", 60, "multiple of 8"),
    ("height", 36, "multiple of 8"),
    ("width", 0, "multiple of 8"),
    ("bitdepth", 9, "bit depth"),
    ("bitdepth", 16, "bit depth"),
    ("layout", 4, "layout must be"),
    ("layout", -1, "layout must be"),
    ("sb128", 2, "sb128"),
    ("width", 16392, "at most 16384"),
    ("height", 32768, "at most 16384"),
])
def test_frame_create_rejects(ctx, field, value, msg):
    sf = R.SynthFrame(64, 64, 10, 0, seed=5, tools=R.TOOL_ALL)
    p = _params(sf)
    setattr(p, field, value)
    with pytest.raises(R.Rav2dError, match=msg):
        R.GpuFrame(ctx, p)
    sf.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["too_many_points", "negative_points"])
def test_grain_data_rejected_and_frame_stays_usable(ctx, case):
    sf = R.SynthFrame(128, 64, 10, 0, seed=9, tools=R.TOOL_ALL)
    good = _params(sf)
    f = R.GpuFrame.from_synth(ctx, sf)
    bad = _params(sf)
    g = bad.grain
    if case == "too_many_points":
        g.num_points[0] = 15
    else:
        g.num_points[1] = -1
    with pytest.raises(R.Rav2dError, match="point"):
        f.set_params(bad)
        f.filter()
        f.apply_grain()
    # the frame still filters and grains exactly like a fresh one
    f.set_params(good)
    f.filter()
    f.apply_grain()
    fresh = R.GpuFrame.from_synth(ctx, sf)
    fresh.filter()
    fresh.apply_grain()
    ctx.synchronize()
    assert f.checksum(1) == fresh.checksum(1)
    assert f.checksum(2) == fresh.checksum(2)
    f.close()
    fresh.close()
    sf.close()


@pytest.mark.gpu
@pytest.mark.parametrize("bitdepth", [8, 10, 12])
@pytest.mark.parametrize("xs", [(10, 200, 100), (60, 60, 180), (200, 120, 40), (90, 30, 150, 240)])
def test_grain_points_that_go_back_match_the_oracle(ctx, bitdepth, xs):
    # the reference decodes such streams: segments with dx <= 0 are skipped
    # and later segments and the tail fill overwrite earlier entries
    import oracle_lib as O
    from test_parity_deblock import _compare
    sf = R.SynthFrame(256, 128, bitdepth, 0, seed=12, tools=R.TOOL_GRAIN)
    g = sf.params.grain
    g.num_points[0] = len(xs)
    for i, x in enumerate(xs):
        g.points[0][i][0], g.points[0][i][1] = x, 40 + 50 * i
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        f.filter()
        f.apply_grain()
        ref = O.filter_frame(sf)
        _compare(O.apply_grain(sf, ref), f.download(2), "points %s" % (xs,))
    finally:
        f.close()
        sf.close()


@pytest.mark.gpu
def test_set_params_rebuilds_the_grain_setup(ctx):
    # filter with parameters A (its grain set-up starts), no apply_grain, then
    # parameters B with another seed and other points: the display planes must
    # be B's, not grain built from A's set-up
    import oracle_lib as O
    from test_parity_deblock import _compare
    sf = R.SynthFrame(352, 288, 10, 0, seed=14, tools=R.TOOL_ALL)
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        f.filter()
        b = _params(sf)
        b.grain_seed = (sf.params.grain_seed + 4321) & 0xffff
        for i in range(b.grain.num_points[0]):
            b.grain.points[0][i][1] = (b.grain.points[0][i][1] * 3 + 17) & 0xff
        f.set_params(b)
        f.filter()
        f.apply_grain()
        ref = O.filter_frame(sf, params=b)
        _compare(ref, f.download(1), "params B output")
        _compare(O.apply_grain(sf, ref, params=b), f.download(2), "params B display")
    finally:
        f.close()
        sf.close()
