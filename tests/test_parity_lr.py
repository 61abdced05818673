"""GPU loop restoration (NS-Wiener, PC-Wiener, GDF) vs the CPU oracle, bit-exact,
alone and as the last stage of the whole in-loop chain."""
import pytest

import rav2d_amd as R
import oracle_lib as O
from test_parity_deblock import _compare

LR_TOOLS = R.TOOL_WIENER | R.TOOL_GDF
CHAIN = R.TOOL_DEBLOCK | R.TOOL_CDEF | R.TOOL_CCSO | R.TOOL_WIENER | R.TOOL_GDF

CASES = [
    # (w, h, bitdepth, sb128, seed, tools, lossless_pct)
    (352, 288, 8, 0, 21, LR_TOOLS, 0),
    (352, 288, 10, 0, 22, LR_TOOLS, 0),
    (640, 360, 10, 1, 23, LR_TOOLS, 10),
    (640, 360, 8, 0, 24, R.TOOL_WIENER, 0),
    (640, 360, 10, 0, 25, R.TOOL_GDF, 0),
    (1920, 1080, 10, 1, 26, CHAIN, 0),
    (1280, 720, 8, 0, 27, CHAIN, 5),
    (136, 72, 10, 0, 28, CHAIN, 0),
    (3840, 2160, 10, 1, 29, CHAIN, 0),
]


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d_%db_sb%d_s%d_t%d_ll%d" % (
    c[0], c[1], c[2], 64 << c[3], c[4], c[5], c[6]))
def test_loop_restoration(ctx, case):
    w, h, bd, sb128, seed, tools, ll = case
    sf = R.SynthFrame(w, h, bd, sb128, seed, tools=tools, lossless_pct=ll)
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        ref = O.filter_frame(sf, R.INLOOP_ALL)
        f.run_pass(R.PASS_LR)
        _compare(ref, f.download(1), "in-loop chain up to LR")
    finally:
        f.close()


@pytest.mark.gpu
@pytest.mark.parametrize("unit_log2", [6, 7, 8, 9])
@pytest.mark.parametrize("ds", [0, 1, 2])
def test_lr_unit_sizes_and_chroma_downsampling(ctx, unit_log2, ds):
    sb128 = 1 if unit_log2 > 6 else 0
    sf = R.SynthFrame(720, 360, 10, sb128, seed=40 + unit_log2 * 3 + ds, tools=LR_TOOLS)
    p = sf.params
    p.lr_unit_size[0] = unit_log2
    p.lr_unit_size[1] = max(5, unit_log2 - 1)
    p.cfl_ds_filter_index = ds
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        ref = O.filter_frame(sf, R.INLOOP_ALL)
        f.run_pass(R.PASS_LR)
        _compare(ref, f.download(1), "LR unit %d ds %d" % (unit_log2, ds))
    finally:
        f.close()


@pytest.mark.gpu
@pytest.mark.parametrize("ref_dst", [0, 1, 3, 5])
@pytest.mark.parametrize("qp", [0, 5])
def test_gdf_tables(ctx, ref_dst, qp):
    sf = R.SynthFrame(512, 256, 10, 0, seed=60 + ref_dst * 7 + qp, tools=R.TOOL_GDF, pc_gdf_pct=100)
    p = sf.params
    p.gdf_ref_dst_idx = ref_dst
    p.gdf_qp_idx = qp
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        ref = O.filter_frame(sf, R.INLOOP_ALL)
        f.run_pass(R.PASS_LR)
        _compare(ref, f.download(1), "GDF ref %d qp %d" % (ref_dst, qp))
    finally:
        f.close()
