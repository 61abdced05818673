"""Argument checks of the C ABI: invalid frame parameters and film-grain data
are rejected on the host with RAV2D_EINVAL and a message, before anything is
launched, and a rejected call leaves the frame usable.

The kernels rely on these checks: planes are multiples of 4 samples wide and
tall (frames are multiples of 8), samples are 8- or 10-bit, and the scaling
points of each plane are at most 14 and non-decreasing (the LUT builder's
segments).  Mirrors the parameter validation of the reference's frame header
and film-grain parsing (rav2d's Rav2dFilmGrainData carries at most 14 points
per plane, filmgrain.rs:22-113)."""
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
    ("bitdepth", 12, "bit depth"),
    ("layout", 3, "4:2:0"),
    ("sb128", 2, "sb128"),
])
def test_frame_create_rejects(ctx, field, value, msg):
    sf = R.SynthFrame(64, 64, 10, 0, seed=5, tools=R.TOOL_ALL)
    p = _params(sf)
    setattr(p, field, value)
    with pytest.raises(R.Rav2dError, match=msg):
        R.GpuFrame(ctx, p)
    sf.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["too_many_points", "negative_points", "decreasing_points"])
def test_grain_data_rejected_and_frame_stays_usable(ctx, case):
    sf = R.SynthFrame(128, 64, 10, 0, seed=9, tools=R.TOOL_ALL)
    good = _params(sf)
    f = R.GpuFrame.from_synth(ctx, sf)
    bad = _params(sf)
    g = bad.grain
    if case == "too_many_points":
        g.num_points[0] = 15
    elif case == "negative_points":
        g.num_points[1] = -1
    else:
        g.num_points[0] = 3
        g.points[0][0][0], g.points[0][1][0], g.points[0][2][0] = 10, 200, 100
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
