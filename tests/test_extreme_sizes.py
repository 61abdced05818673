"""Frame shapes at the edges of the accepted range against the oracle: the
smallest frames (8x8: 4x4 chroma planes, one restoration block, one grain
block) and strips 16384 samples long in either direction (the largest
dimension the library accepts: work-item coordinates near the int16 range,
256 superblock columns or 512 tile rows in one launch)."""
import pytest

import rav2d_amd as R
import oracle_lib as O
from test_parity_deblock import _compare

# (w, h, bitdepth, sb128, mask_flags, monochrome)
CASES = [
    (8, 8, 10, 0, 0, False),
    (16, 8, 8, 1, R.MASK_TXPART | R.MASK_SUBPU | R.MASK_SEG, False),
    (8, 16, 12, 0, R.MASK_TXPART | R.MASK_SUBPU, False),
    (24, 40, 10, 1, 0, True),
    (16384, 64, 10, 1, 0, False),
    (64, 16384, 10, 0, R.MASK_TXPART | R.MASK_SUBPU, False),
    (16384, 16, 8, 0, 0, False),
    (8, 4096, 12, 1, 0, False),
]


def _frame(c):
    w, h, bd, sb, mf, mono = c
    return R.SynthFrame(w, h, bd, sb, 99 + w + h, mask_flags=mf, lossless_pct=5 if mf & R.MASK_SEG else 0,
                        monochrome=mono)


def _tag(c):
    return "%dx%d_%db_sb%d_m%d%s" % (c[0], c[1], c[2], 64 << c[3], c[4], "_i400" if c[5] else "")


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


@pytest.mark.parametrize("c", CASES[:4], ids=_tag)
def test_oracle_filters_small_frames(c):
    """The oracle runs the whole chain on the smallest frames (every filter
    changes samples; nothing is read outside the planes, as the ASan build of
    the oracle confirms)."""
    sf = _frame(c)
    try:
        ref = O.filter_frame(sf)
        assert (ref[0] != sf.planes()[0]).mean() > 0.5
        O.apply_grain(sf, ref)
    finally:
        sf.close()


@pytest.mark.gpu
@pytest.mark.parametrize("c", CASES, ids=_tag)
def test_extreme_size_matches_oracle(ctx, c):
    sf = _frame(c)
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        f.filter()
        f.apply_grain()
        ref = O.filter_frame(sf)
        _compare(ref, f.download(1), _tag(c) + " output")
        _compare(O.apply_grain(sf, ref), f.download(2), _tag(c) + " display")
    finally:
        f.close()
        sf.close()


@pytest.mark.gpu
def test_extreme_sizes_in_one_batch(ctx):
    sfs = [_frame(c) for c in CASES]
    frames = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs]
    b = R.GpuBatch(ctx, frames)
    try:
        b.run()
        ctx.synchronize()
        for c, sf, f in zip(CASES, sfs, frames):
            ref = O.filter_frame(sf)
            _compare(ref, f.download(1), "batch " + _tag(c))
            _compare(O.apply_grain(sf, ref), f.download(2), "batch " + _tag(c) + " display")
    finally:
        b.close()
        for f in frames:
            f.close()
        for sf in sfs:
            sf.close()
