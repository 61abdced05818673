"""rav2d_gpu_apply_grain_planes_noalloc (static scratch, allocates nothing)
gives the planes rav2d_gpu_apply_grain_planes gives, on two calls in a row,
the second on a smaller picture (the static scratch keeps the first call's
contents), and rejects pictures beyond its maximum."""
import ctypes as C

import numpy as np
import pytest

import rav2d_amd as R


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


def _grain(ctx, sf, noalloc):
    w, h, bd = sf.width, sf.height, sf.bitdepth
    src = R.Picture(ctx, w, h, bd)
    dst = R.Picture(ctx, w, h, bd)
    try:
        src.write(sf.planes())
        fn = ctx.lib.rav2d_gpu_apply_grain_planes_noalloc if noalloc else ctx.lib.rav2d_gpu_apply_grain_planes
        R._check(fn(ctx.h, C.byref(src.planes), C.byref(dst.planes), w, h, bd, 1, C.byref(sf.params.grain),
                    sf.params.grain_seed, None), "apply_grain_planes%s" % ("_noalloc" if noalloc else ""))
        ctx.synchronize()
        return dst.read()
    finally:
        src.close()
        dst.close()


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [10, 8])
def test_noalloc_equals_allocating_entry_twice_second_smaller(ctx, bd):
    for w, h, seed in [(1920, 1080, 91 + bd), (352, 288, 92 + bd)]:
        sf = R.SynthFrame(w, h, bd, 0, seed, tools=R.TOOL_GRAIN)
        try:
            want = _grain(ctx, sf, False)
            got = _grain(ctx, sf, True)
            for pl in range(3):
                assert np.array_equal(want[pl], got[pl]), (w, h, pl)
            assert not np.array_equal(want[0], sf.planes()[0])  # grain was applied
        finally:
            sf.close()


@pytest.mark.gpu
def test_noalloc_rejects_pictures_beyond_its_scratch(ctx):
    src = R.Picture(ctx, 8200, 64, 8)
    try:
        g = R.FilmGrainData()
        g.num_points[0] = 1
        g.scaling_shift, g.ar_coeff_shift = 8, 6
        with pytest.raises(R.Rav2dError, match="8192"):
            R._check(ctx.lib.rav2d_gpu_apply_grain_planes_noalloc(ctx.h, C.byref(src.planes), C.byref(src.planes),
                                                                  8200, 64, 8, 1, C.byref(g), 1, None), "noalloc")
    finally:
        src.close()
