"""12-bit frames (the reference's u16 BitDepth16 path with bitdepth 12,
pixel.rs:120-144): deblock thresholds with hbd = 2 (lf_mask.rs:332-342),
CDEF/CCSO with bitdepth_min_8 = 4, loop restoration with GDF's 12-bit samples
shifted into the 10-bit domain and no rounding shift in gdf_add
(looprestoration.rs:273-329, 1450-1453), film grain with 4096-entry scaling
LUTs (filmgrain.rs:58-113).  The GPU chain equals the oracle through the C ABI,
per pass, whole, with grain, and batched with 8- and 10-bit frames."""
import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O
from test_parity_deblock import _compare

ORACLE_DEBLOCK_COLS_ONLY = 1 << 8


def test_thresholds_12bit_closed_form():
    lib = O.load()
    # quant: dq_lookup clamps qidx to 0..255 + 48 * 2
    assert lib.oracle_deblock_quant_thr(2, 400) == lib.oracle_deblock_quant_thr(2, 351)
    assert lib.oracle_deblock_quant_thr(2, 351) == lib.oracle_deblock_quant_thr(2, 255)
    # side: q_ind = clip(qidx - 24 * 4, 0, 295), max(T + (16 >> 4), 0) >> (5 - 4);
    # DEBLOCK_SIDE_THRESHOLDS[0] = -16, [295] = 1678
    assert lib.oracle_deblock_side_thr(2, 0) == 0
    assert lib.oracle_deblock_side_thr(2, 96) == 0
    assert lib.oracle_deblock_side_thr(2, 295 + 96) == (1678 + 1) >> 1
    assert lib.oracle_deblock_side_thr(2, 500) == (1678 + 1) >> 1
    v10 = [lib.oracle_deblock_side_thr(1, q) for q in range(48, 344)]
    v12 = [lib.oracle_deblock_side_thr(2, q) for q in range(96, 392)]
    # 10 bit: (T + 4) >> 3; 12 bit: (T + 1) >> 1: the 12-bit threshold is about 4x the 10-bit one
    assert all(abs(b - 4 * a) <= 4 for a, b in zip(v10, v12))


def test_oracle_12bit_uniform_frame_unchanged():
    sf = R.SynthFrame(192, 128, 12, 0, seed=71, tools=R.TOOL_DEBLOCK | R.TOOL_CDEF | R.TOOL_WIENER)
    try:
        from test_oracle_kat import _fill
        v = [2000, 1200, 3000]
        _fill(sf, v)
        out = O.filter_frame(sf)
        for pl in range(3):
            assert (out[pl] == v[pl]).all()
    finally:
        sf.close()


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


CASES = [
    # (w, h, sb128, seed, tools, mask_flags, lossless_pct)
    (352, 288, 0, 81, R.TOOL_ALL, 0, 0),
    (640, 360, 1, 82, R.TOOL_ALL, R.MASK_TXPART | R.MASK_SUBPU, 5),
    (264, 136, 0, 83, R.TOOL_ALL & ~R.TOOL_GRAIN, R.MASK_SEG, 10),
    (512, 256, 0, 84, R.TOOL_GDF | R.TOOL_WIENER, 0, 0),
    (1920, 1080, 1, 85, R.TOOL_ALL, 0, 0),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d_sb%d_s%d_t%d_m%d" % (c[0], c[1], 64 << c[2], c[3], c[4],
                                                                                c[5]))
def test_12bit_chain_matches_oracle(ctx, case):
    w, h, sb128, seed, tools, flags, ll = case
    sf = R.SynthFrame(w, h, 12, sb128, seed, tools=tools, mask_flags=flags, lossless_pct=ll,
                      pc_gdf_pct=100 if tools == R.TOOL_GDF | R.TOOL_WIENER else 50)
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        stages = [(R.PASS_DEBLOCK_V, R.INLOOP_DEBLOCK | ORACLE_DEBLOCK_COLS_ONLY),
                  (R.PASS_DEBLOCK_H, R.INLOOP_DEBLOCK),
                  (R.PASS_CDEF, R.INLOOP_DEBLOCK | R.INLOOP_CDEF | R.INLOOP_CCSO)]
        for up, mask in stages:
            f.run_pass(up)
            _compare(O.filter_frame(sf, mask), f.download(1), "12-bit pass %d" % up)
        f.filter()
        out = f.download(1)
        ref = O.filter_frame(sf)
        _compare(ref, out, "12-bit chain")
        assert max(int(a.max()) for a in out) > 1023
        if sf.params.grain_present:
            f.apply_grain()
            _compare(O.apply_grain(sf, ref), f.download(2), "12-bit grain")
    finally:
        f.close()
        sf.close()


# batch configurations: (mixed_depths, deblock_rows, grid, xcd_remap); mixed_depths 1 runs
# one launch per pass over the three bit depths
BATCH_CONFIGS = [(0, 0, 0, 0), (1, 0, 0, 0), (1, 64, 0, 1), (1, 32, 40, 1), (0, 64, 24, 0)]


@pytest.mark.gpu
@pytest.mark.parametrize("conf", BATCH_CONFIGS, ids=lambda c: "mixed%d_dbrows%d_grid%d_xcd%d" % c)
def test_batch_mixes_8_10_and_12bit(ctx, conf):
    spec = [(640, 360, 8, 0, 91), (352, 288, 12, 1, 92), (640, 360, 10, 1, 93), (264, 136, 12, 0, 94),
            (1280, 720, 12, 0, 95)]
    sfs = [R.SynthFrame(w, h, bd, sb, s) for w, h, bd, sb, s in spec]
    frames = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs]
    cfg = R.default_batch_config()
    cfg.mixed_depths, cfg.deblock_rows = conf[0], conf[1]
    for p in range(R.NUM_PASSES):
        cfg.grid[p], cfg.xcd_remap[p] = conf[2], conf[3]
    b = R.GpuBatch(ctx, frames, cfg)
    try:
        for _ in range(2):
            b.run()
        ctx.synchronize()
        for sf, f in zip(sfs, frames):
            ref = O.filter_frame(sf)
            _compare(ref, f.download(1), "batch %d-bit" % sf.bitdepth)
            if sf.params.grain_present:
                _compare(O.apply_grain(sf, ref), f.download(2), "batch grain %d-bit" % sf.bitdepth)
    finally:
        b.close()
        for f in frames:
            f.close()
        for sf in sfs:
            sf.close()
