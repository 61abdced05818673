"""GPU film grain on the output copy vs the CPU oracle, bit-exact."""
import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O
from test_parity_deblock import _compare

CASES = [
    # (w, h, bitdepth, sb128, seed)
    (352, 288, 8, 0, 31),
    (352, 288, 10, 0, 32),
    (1920, 1080, 10, 0, 33),
    (640, 360, 8, 0, 34),
    (136, 72, 10, 0, 35),
    # the benchmark's roofline workload: every tool, 128x128 superblocks
    (3840, 2160, 10, 1, 36),
    (2560, 1440, 10, 0, 37),
]

# sizes of the mixed batch the grain-free (8-bit) frames take
CASES_NO_GRAIN = [
    # (w, h, bitdepth, sb128, seed)
    (2560, 1440, 8, 1, 38),
    (3840, 2160, 8, 0, 39),
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
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d_%db_sb%d_s%d" % (c[0], c[1], c[2], 64 << c[3], c[4]))
def test_grain_chain(ctx, case):
    w, h, bd, sb128, seed = case
    sf = R.SynthFrame(w, h, bd, sb128, seed, tools=R.TOOL_ALL)
    _run(ctx, sf, "grain")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES_NO_GRAIN,
                         ids=lambda c: "%dx%d_%db_sb%d_s%d" % (c[0], c[1], c[2], 64 << c[3], c[4]))
def test_chain_without_grain(ctx, case):
    # the display copy of a frame without grain equals its in-loop output
    w, h, bd, sb128, seed = case
    sf = R.SynthFrame(w, h, bd, sb128, seed, tools=R.TOOL_ALL & ~R.TOOL_GRAIN)
    _run(ctx, sf, "no grain")


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
