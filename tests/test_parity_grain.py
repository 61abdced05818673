"""GPU film grain on the output copy vs the CPU oracle, bit-exact."""
import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O
from test_parity_deblock import _compare

CASES = [
    # (w, h, bitdepth, seed)
    (352, 288, 8, 31),
    (352, 288, 10, 32),
    (1920, 1080, 10, 33),
    (640, 360, 8, 34),
    (136, 72, 10, 35),
]


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


def _run(ctx, sf, tag):
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        f.filter()
        f.apply_grain()
        ref_out = O.filter_frame(sf)
        _compare(ref_out, f.download(1), tag + " (in-loop output)")
        ref_disp = O.apply_grain(sf, ref_out)
        _compare(ref_disp, f.download(2), tag + " (display copy)")
    finally:
        f.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d_%db_s%d" % c)
def test_grain_chain(ctx, case):
    w, h, bd, seed = case
    sf = R.SynthFrame(w, h, bd, 0, seed, tools=R.TOOL_ALL)
    _run(ctx, sf, "grain")


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["no_overlap_bs16", "overlap_bs32", "restricted_mcid", "uv_points", "lag0",
                                     "lag1", "lag2", "no_luma_points"])
def test_grain_parameter_variants(ctx, variant):
    sf = R.SynthFrame(512, 272, 10, 0, seed=77, tools=R.TOOL_GRAIN)
    g = sf.params.grain
    if variant == "no_overlap_bs16":
        g.overlap_flag, g.block_size = 0, 0
    elif variant == "overlap_bs32":
        g.overlap_flag, g.block_size = 1, 1
    elif variant == "restricted_mcid":
        g.clip_to_restricted_range, g.mc_identity = 1, 1
    elif variant == "uv_points":
        g.chroma_scaling_from_luma = 0
        for pl in (1, 2):
            g.num_points[pl] = 4
            for i in range(4):
                g.points[pl][i][0] = 20 + 60 * i
                g.points[pl][i][1] = 40 + 30 * i + pl
        g.uv_mult[0], g.uv_luma_mult[0], g.uv_offset[0] = 64, -30, 17
        g.uv_mult[1], g.uv_luma_mult[1], g.uv_offset[1] = -100, 120, -200
    elif variant == "lag0":
        g.ar_coeff_lag = 0
    elif variant in ("lag1", "lag2"):
        # the AR window's taps beyond the lag are zero; chroma keeps its luma term
        g.ar_coeff_lag = int(variant[-1])
    elif variant == "no_luma_points":
        g.num_points[0] = 0
        g.clip_to_restricted_range = 1
    _run(ctx, sf, variant)
