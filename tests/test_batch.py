"""Batched launches (one launch per pass over many frames of mixed size, bit
depth and tool set) against the CPU oracle, bit-exact, for every launch
shape: one workgroup per item, persistent grids, XCD-contiguous ranges,
graph replay and eager launches."""
import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O
from test_parity_deblock import _compare

T = R
# (w, h, bitdepth, sb128, seed, tools)
FRAMES = [
    (640, 360, 10, 1, 101, T.TOOL_ALL),
    (352, 288, 8, 0, 102, T.TOOL_ALL & ~T.TOOL_GRAIN),
    (136, 72, 10, 0, 103, T.TOOL_ALL),
    (256, 144, 8, 1, 104, T.TOOL_DEBLOCK),
    (320, 184, 10, 0, 105, T.TOOL_CDEF | T.TOOL_CCSO | T.TOOL_GRAIN),
    (200, 120, 10, 1, 106, T.TOOL_WIENER | T.TOOL_GDF),
    (64, 64, 8, 0, 107, 0),
    (480, 272, 10, 0, 108, T.TOOL_GRAIN),
]


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def synth():
    sfs = [R.SynthFrame(w, h, bd, sb, seed, tools=tools) for (w, h, bd, sb, seed, tools) in FRAMES]
    refs = []
    for sf in sfs:
        out = O.filter_frame(sf)
        refs.append((out, O.apply_grain(sf, out)))
    yield sfs, refs
    for sf in sfs:
        sf.close()


def _check_all(frames, refs, tag):
    for i, (f, (ref_out, ref_disp)) in enumerate(zip(frames, refs)):
        _compare(ref_out, f.download(1), "%s frame %d %s (output)" % (tag, i, FRAMES[i][:3]))
        _compare(ref_disp, f.download(2), "%s frame %d %s (display)" % (tag, i, FRAMES[i][:3]))


def _config(use_graph, grid, xcd, deblock_rows=0, mixed=0):
    cfg = R.default_batch_config()
    cfg.use_graph = use_graph
    for p in range(R.NUM_PASSES):
        cfg.grid[p] = grid
        cfg.xcd_remap[p] = xcd
    cfg.deblock_rows = deblock_rows
    cfg.mixed_depths = mixed
    return cfg


# (use_graph, grid, xcd_remap, deblock_rows, mixed_depths): deblock_rows 32 / 64
# runs both deblocking passes as one tile pass; mixed_depths 1 one launch per
# pass over the 8- and 10-bit frames
SHAPES = [(1, 0, 0, 0, 0), (1, 0, 1, 0, 0), (1, 64, 0, 0, 0), (1, 24, 1, 0, 0), (0, 0, 0, 0, 0), (0, 7, 1, 0, 0),
          (1, 0, 0, 32, 0), (1, 0, 1, 64, 0), (0, 24, 1, 32, 0), (1, 64, 0, 64, 0),
          (1, 0, 0, 0, 1), (1, 0, 1, 64, 1), (0, 24, 1, 32, 1), (1, 7, 0, 0, 1)]


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "graph%d_grid%d_xcd%d_dbrows%d_mixed%d" % s)
def test_batch_matches_oracle(ctx, synth, shape):
    sfs, refs = synth
    frames = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs]
    b = R.GpuBatch(ctx, frames, _config(*shape))
    try:
        b.run()
        ctx.synchronize()
        _check_all(frames, refs, "run 1")
        # a replay reads the untouched input planes again and writes the same
        b.run()
        b.run()
        ctx.synchronize()
        _check_all(frames, refs, "run 3")
    finally:
        b.close()
        for f in frames:
            f.close()


@pytest.mark.gpu
def test_batch_replans_after_param_change(ctx, synth):
    sfs, refs = synth
    frames = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs]
    b = R.GpuBatch(ctx, frames)
    try:
        b.run()
        ctx.synchronize()
        # switch CDEF/CCSO off on frame 0: the next run must re-plan that frame
        p = R.FrameParams.from_buffer_copy(sfs[0].params)
        p.inloop_filters &= ~(R.INLOOP_CDEF | R.INLOOP_CCSO)
        frames[0].set_params(p)
        b.run()
        ctx.synchronize()
        ref = O.filter_frame(sfs[0], params=p)
        _compare(ref, frames[0].download(1), "re-planned frame 0 (output)")
        _compare(O.apply_grain(sfs[0], ref, params=p), frames[0].download(2), "re-planned frame 0 (display)")
        _check_all(frames[1:], refs[1:], "others")
    finally:
        b.close()
        for f in frames:
            f.close()


@pytest.mark.gpu
def test_batch_autotune_keeps_results(ctx, synth):
    sfs, refs = synth
    frames = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs]
    b = R.GpuBatch(ctx, frames)
    try:
        cfg = b.autotune(iters=2)
        assert cfg.use_graph == 1
        b.run()
        ctx.synchronize()
        _check_all(frames, refs, "autotuned")
        items, nbytes, nf = b.info(R.PASS_LR)
        assert nf == len(frames) and items > 0 and nbytes > 0
        assert b.pixels() == sum(w * h for (w, h, *_rest) in FRAMES)
    finally:
        b.close()
        for f in frames:
            f.close()


@pytest.mark.gpu
def test_batch_without_grain_leaves_display_alone(ctx, synth):
    sfs, refs = synth
    frames = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs[:3]]
    cfg = R.default_batch_config()
    cfg.apply_grain = 0
    b = R.GpuBatch(ctx, frames, cfg)
    try:
        b.run()
        ctx.synchronize()
        for i, f in enumerate(frames):
            _compare(refs[i][0], f.download(1), "no-grain batch frame %d" % i)
        assert b.info(R.PASS_GRAIN)[0] == 0
    finally:
        b.close()
        for f in frames:
            f.close()


@pytest.mark.gpu
def test_batch_switches_deblocking_form(ctx, synth):
    """set_config between separate and fused deblocking re-plans the batch;
    every form gives the oracle's planes, and tile rows other than 32 / 64
    are rejected with the configuration unchanged."""
    sfs, refs = synth
    frames = [R.GpuFrame.from_synth(ctx, sf) for sf in sfs]
    b = R.GpuBatch(ctx, frames)
    try:
        for rows in (64, 0, 32):
            cfg = b.config()
            cfg.deblock_rows = rows
            b.set_config(cfg)
            b.run()
            ctx.synchronize()
            _check_all(frames, refs, "deblock_rows %d" % rows)
        bad = b.config()
        bad.deblock_rows = 48
        with pytest.raises(R.Rav2dError):
            b.set_config(bad)
        assert b.config().deblock_rows == 32
        with pytest.raises(R.Rav2dError):
            R.GpuBatch(ctx, frames[:1], bad)
    finally:
        b.close()
        for f in frames:
            f.close()
