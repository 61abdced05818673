"""Seeded random sweep against the oracle: frames of random size, bit depth,
superblock size, tool subset, mask features (transform partitions, sub-PU
and TIP masks, segmentation), lossless / skip / restoration percentages and
layout (4:2:0, 4:2:2, 4:4:4 or I400), each filtered and grained on the GPU per frame, and
all of them together in batches (one launch per pass over every frame)."""
import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O
from test_parity_deblock import _compare
from test_layouts_422_444 import _y_stride

SWEEP = 40


def _cases():
    rng = np.random.default_rng(20261204)
    out = []
    for i in range(SWEEP):
        out.append(dict(
            seed=7000 + i,
            w=int(rng.integers(6, 61)) * 8,   # 48 .. 480
            h=int(rng.integers(5, 41)) * 8,   # 40 .. 320
            bd=int(rng.choice([8, 10, 12])),
            sb=int(rng.integers(0, 2)),
            tools=int(rng.integers(1, 64)),   # a non-empty subset of the six tools
            mask_flags=int(rng.integers(0, 8)),
            ll=int(rng.choice([0, 0, 5, 20])),
            skip=int(rng.integers(0, 80)),
            pcg=int(rng.integers(0, 101)),
            mono=bool(rng.random() < 0.15),
            layout=int(rng.choice([R.PIXEL_LAYOUT_I420, R.PIXEL_LAYOUT_I420, R.PIXEL_LAYOUT_I422,
                                   R.PIXEL_LAYOUT_I444])),
        ))
    return out


CASES = _cases()


def _frame(c):
    return R.SynthFrame(c["w"], c["h"], c["bd"], c["sb"], c["seed"], tools=c["tools"], lossless_pct=c["ll"],
                        skip_pct=c["skip"], pc_gdf_pct=c["pcg"], mask_flags=c["mask_flags"],
                        layout=R.PIXEL_LAYOUT_I400 if c["mono"] else c["layout"])


def _tag(c):
    lay = "i400" if c["mono"] else {R.PIXEL_LAYOUT_I420: "i420", R.PIXEL_LAYOUT_I422: "i422",
                                     R.PIXEL_LAYOUT_I444: "i444"}[c["layout"]]
    return "%dx%d_%db_sb%d_t%d_m%d_ll%d_s%d_%s" % (c["w"], c["h"], c["bd"], 64 << c["sb"], c["tools"], c["mask_flags"],
                                                  c["ll"], c["seed"], lay)


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


def test_sweep_covers_the_space():
    """The drawn cases reach every bit depth, both superblock sizes, every
    mask feature, lossless frames, monochrome frames and each tool both on
    and off."""
    assert {c["bd"] for c in CASES} == {8, 10, 12}
    assert {c["sb"] for c in CASES} == {0, 1}
    for bit in (R.MASK_TXPART, R.MASK_SUBPU, R.MASK_SEG):
        assert any(c["mask_flags"] & bit for c in CASES)
    for bit in (R.TOOL_DEBLOCK, R.TOOL_CDEF, R.TOOL_CCSO, R.TOOL_WIENER, R.TOOL_GDF, R.TOOL_GRAIN):
        assert any(c["tools"] & bit for c in CASES) and any(not c["tools"] & bit for c in CASES)
    assert any(c["ll"] for c in CASES) and any(c["mono"] for c in CASES)
    assert {c["layout"] for c in CASES if not c["mono"]} == {R.PIXEL_LAYOUT_I420, R.PIXEL_LAYOUT_I422,
                                                             R.PIXEL_LAYOUT_I444}


@pytest.mark.gpu
@pytest.mark.parametrize("c", CASES, ids=_tag)
def test_sweep_frame_matches_oracle(ctx, c):
    sf = _frame(c)
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        f.filter()
        f.apply_grain()
        ref = O.filter_frame(sf, y_stride=_y_stride(f))
        _compare(ref, f.download(1), _tag(c) + " output")
        _compare(O.apply_grain(sf, ref), f.download(2), _tag(c) + " display")
    finally:
        f.close()
        sf.close()


@pytest.mark.gpu
@pytest.mark.parametrize("conf", [(0, 0), (64, 0), (32, 1)], ids=lambda c: "dbrows%d_mixed%d" % c)
def test_sweep_batch_matches_oracle(ctx, conf):
    sfs = [_frame(c) for c in CASES]
    frames = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs]
    cfg = R.default_batch_config()
    cfg.deblock_rows, cfg.mixed_depths = conf
    b = R.GpuBatch(ctx, frames, cfg)
    try:
        b.run()
        b.run()
        ctx.synchronize()
        for c, sf, f in zip(CASES, sfs, frames):
            ref = O.filter_frame(sf, y_stride=_y_stride(f))
            _compare(ref, f.download(1), "batch " + _tag(c))
            _compare(O.apply_grain(sf, ref), f.download(2), "batch " + _tag(c) + " display")
    finally:
        b.close()
        for f in frames:
            f.close()
        for sf in sfs:
            sf.close()
