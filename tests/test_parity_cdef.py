"""GPU CDEF + CCSO (after deblocking) vs the CPU oracle, bit-exact."""
import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O
from test_parity_deblock import _compare

CASES = [
    # (w, h, bitdepth, sb128, seed, tools)
    (352, 288, 8, 0, 11, R.TOOL_DEBLOCK | R.TOOL_CDEF | R.TOOL_CCSO),
    (352, 288, 10, 0, 12, R.TOOL_DEBLOCK | R.TOOL_CDEF | R.TOOL_CCSO),
    (640, 360, 10, 1, 13, R.TOOL_DEBLOCK | R.TOOL_CDEF | R.TOOL_CCSO),
    (1920, 1080, 10, 1, 14, R.TOOL_DEBLOCK | R.TOOL_CDEF | R.TOOL_CCSO),
    (1280, 720, 8, 0, 15, R.TOOL_CDEF),               # CDEF alone
    (640, 360, 8, 0, 16, R.TOOL_DEBLOCK | R.TOOL_CCSO),  # CCSO without CDEF
    (136, 72, 10, 0, 17, R.TOOL_DEBLOCK | R.TOOL_CDEF | R.TOOL_CCSO),  # ragged
]


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d_%db_sb%d_s%d_t%d" % (c[0], c[1], c[2], 64 << c[3], c[4], c[5]))
def test_cdef_ccso(ctx, case):
    w, h, bd, sb128, seed, tools = case
    sf = R.SynthFrame(w, h, bd, sb128, seed, tools=tools)
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        ref = O.filter_frame(sf, R.INLOOP_DEBLOCK | R.INLOOP_CDEF | R.INLOOP_CCSO)
        f.run_pass(R.PASS_CDEF)
        _compare(ref, f.download(1), "deblock+cdef+ccso")
    finally:
        f.close()


@pytest.mark.gpu
def test_cdef_on_skiptx_and_all_strength_codes(ctx):
    """Every strength code 0..63 for luma and chroma, every damping, on_skiptx both ways."""
    for damping in (3, 6):
        for on_skip in (0, 1):
            sf = R.SynthFrame(512, 256, 10, 0, seed=100 + damping * 2 + on_skip, tools=R.TOOL_CDEF)
            p = sf.params
            p.cdef_damping = damping
            p.cdef_on_skiptx = on_skip
            p.cdef_n_strengths = 8
            for i in range(8):
                p.cdef_y_strength[i] = (damping * 8 + i * 9 + on_skip) % 64
                p.cdef_uv_strength[i] = (63 - i * 7 - damping) % 64
            f = R.GpuFrame.from_synth(ctx, sf)
            try:
                ref = O.filter_frame(sf, R.INLOOP_CDEF)
                f.run_pass(R.PASS_CDEF)
                _compare(ref, f.download(1), "cdef d%d skip%d" % (damping, on_skip))
            finally:
                f.close()


# offsetof(Rav2dAv2Filter, ccso): filter_y and filter_uv (2 x 64 x 5 x 4 uint16 each),
# qidx[16] uint16, gdf[16], cdef_idx[16]
CCSO_FLAGS_OFF = 2 * (2 * 64 * 5 * 4 * 2) + 16 * 2 + 16 + 16


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [(0, 0, 0, 3), (0, 1, 4, 2), (0, 0, 6, 1), (1, 0, 0, 5)],
                         ids=lambda m: "bo%d_clf%d_ext%d_band%d" % m)
def test_ccso_modes(ctx, mode):
    """CCSO on every plane in most 256x256 regions (interior superblocks and
    frame-edge ones), edge-classified and band-only, several neighbour offsets,
    with non-zero offset tables so that every index maps to a distinct value."""
    bo, clf, ext, band = mode
    sf = R.SynthFrame(576, 320, 10, 0, seed=120 + ext, tools=R.TOOL_DEBLOCK | R.TOOL_CDEF | R.TOOL_CCSO)
    p = sf.params
    for pl in range(3):
        c = p.ccso[pl]
        c.enabled, c.bo_only, c.edge_clf, c.ext_filter_support, c.max_band_log2 = 1, bo, clf, ext, band
        c.scale_idx, c.quant_idx = pl & 3, (pl + 1) & 3
        for k in range(64):
            c.filter_off[k] = (k * 37 + pl * 11 + 5) & 0xFF
    m = sf.mask_bytes()
    for i in range(sf.n_mask):
        for pl in range(3):
            m[i * R.AV2FILTER_SIZE + CCSO_FLAGS_OFF + pl] = 0 if (i + pl) % 4 == 3 else 1
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        ref = O.filter_frame(sf, R.INLOOP_DEBLOCK | R.INLOOP_CDEF | R.INLOOP_CCSO)
        f.run_pass(R.PASS_CDEF)
        _compare(ref, f.download(1), "ccso bo%d clf%d ext%d band%d" % mode)
    finally:
        f.close()
