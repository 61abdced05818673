"""Serving: the continuous-batching request loop returns, for every job, the
planes that the same frame filtered alone gives (tolerance 0: the 64-bit
plane hashes are equal, and equal the hash of the CPU oracle's planes), under
Poisson arrivals from many streams and under a burst; memory admission
follows the context's HBM budget."""
import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O

T = R
# one synthetic source per stream: (w, h, bitdepth, sb128, seed, tools)
STREAMS = [
    (320, 184, 10, 1, 201, T.TOOL_ALL),
    (256, 144, 8, 0, 202, T.TOOL_ALL & ~T.TOOL_GRAIN),
    (136, 72, 10, 0, 203, T.TOOL_ALL),
    (640, 360, 8, 1, 204, T.TOOL_DEBLOCK | T.TOOL_CDEF),
    (200, 120, 10, 0, 205, T.TOOL_WIENER | T.TOOL_GDF | T.TOOL_GRAIN),
    (96, 64, 8, 0, 206, 0),
]
SLOTS = 2


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def sources():
    sfs = [R.SynthFrame(w, h, bd, sb, seed, tools=tools) for (w, h, bd, sb, seed, tools) in STREAMS]
    expect = []
    for sf in sfs:
        out = O.filter_frame(sf)
        expect.append((R.plane_checksum(out), R.plane_checksum(O.apply_grain(sf, out))))
    yield sfs, expect
    for sf in sfs:
        sf.close()


@pytest.mark.gpu
def test_checksum_kernel_matches_host(ctx, sources):
    sfs, expect = sources
    for i, sf in enumerate(sfs[:3]):
        f = R.GpuFrame.from_synth(ctx, sf)
        try:
            assert f.checksum(0) == R.plane_checksum(sf.planes()), "input hash, stream %d" % i
            f.filter()
            f.apply_grain()
            assert (f.checksum(1), f.checksum(2)) == expect[i], "alone hashes vs oracle, stream %d" % i
        finally:
            f.close()


def _frames(ctx, sfs):
    return [[R.GpuFrame.from_synth(ctx, sf) for _ in range(SLOTS)] for sf in sfs]


@pytest.mark.gpu
@pytest.mark.parametrize("max_batch,max_inflight", [(16, 3), (1, 2), (4, 1)])
def test_poisson_mix_equals_alone(ctx, sources, max_batch, max_inflight):
    sfs, expect = sources
    frames = _frames(ctx, sfs)
    cfg = R.default_serve_config()
    cfg.max_batch, cfg.max_inflight, cfg.checksum = max_batch, max_inflight, 1
    srv = R.GpuServer(ctx, cfg)
    try:
        res, js, jl = srv.simulate(frames, SLOTS, fps=1500.0, n_jobs=150, seed=9)
    finally:
        srv.close()
        for fs in frames:
            for f in fs:
                f.close()
    assert len(res) == 150
    sizes = []
    for j, r in enumerate(res):
        assert r.tag == j and r.status == 0, (j, r.status)
        assert r.t_done >= r.t_launch >= r.t_submit - 1e-3
        got = (r.checksum[0], r.checksum[1])
        assert got == expect[js[j]], "job %d (stream %d slot %d, batch of %d) differs from alone" % (
            j, js[j], jl[j], r.batch_frames)
        sizes.append(r.batch_frames)
    assert max(sizes) <= max_batch


@pytest.mark.gpu
def test_burst_is_batched_and_exact(ctx, sources):
    sfs, expect = sources
    frames = _frames(ctx, sfs)
    cfg = R.default_serve_config()
    cfg.checksum = 1
    srv = R.GpuServer(ctx, cfg)
    flat = [(i, f) for i, fs in enumerate(frames) for f in fs]
    try:
        tags = {}
        for k in range(4):  # each frame four times: later jobs of a frame wait for the earlier
            for n, (i, f) in enumerate(flat):
                tag = k * len(flat) + n
                tags[tag] = i
                srv.submit(f, tag)
        srv.drain()
        got = []
        while len(got) < len(tags):
            got += srv.poll(64, 100)
    finally:
        srv.close()
        for fs in frames:
            for f in fs:
                f.close()
    assert sorted(r.tag for r in got) == sorted(tags)
    for r in got:
        assert r.status == 0
        assert (r.checksum[0], r.checksum[1]) == expect[tags[r.tag]], "burst job %d" % r.tag
    assert max(r.batch_frames for r in got) > 1


@pytest.mark.gpu
def test_admission_follows_hbm_budget(sources):
    sfs, _ = sources
    sf = sfs[0]
    budget = 48 << 20
    c = R.GpuContext(0, hbm_budget=budget)
    frames = []
    try:
        with pytest.raises(R.Rav2dError, match="budget"):
            for _ in range(1000):
                frames.append(R.GpuFrame.from_synth(c, sf))
        assert frames and c.bytes_in_use() <= budget
        held = c.bytes_in_use()
        frames.pop().close()  # freeing a frame admits the next one
        assert c.bytes_in_use() < held
        frames.append(R.GpuFrame.from_synth(c, sf))
        srv = R.GpuServer(c)
        try:
            for n, f in enumerate(frames):
                srv.submit(f, n)
            srv.drain()
            res = []
            while len(res) < len(frames):
                res += srv.poll(64, 100)
            assert all(r.status == 0 for r in res)
        finally:
            srv.close()
    finally:
        for f in frames:
            f.close()
        c.close()


@pytest.mark.gpu
def test_injected_allocation_failure_is_reported_and_recovers(ctx, sources):
    # frame creation whose device allocation fails returns ENOMEM, holds no
    # memory, and the next creation succeeds
    sfs, _ = sources
    before = ctx.bytes_in_use()
    R.inject_failure(R.FAIL_ALLOC, 1)
    try:
        with pytest.raises(R.Rav2dError, match="-12"):
            R.GpuFrame(ctx, sfs[0].params)
    finally:
        R.inject_failure(R.FAIL_ALLOC, 0)
    assert ctx.bytes_in_use() == before
    f = R.GpuFrame.from_synth(ctx, sfs[0])
    f.close()


@pytest.mark.gpu
def test_injected_launch_failure_fails_its_batch_only(ctx, sources):
    # a batch whose launch fails reports the error in every job of it; the
    # frames stay usable and later jobs give the planes the frame gives alone
    sfs, expect = sources
    frames = [R.GpuFrame.from_synth(ctx, sfs[i]) for i in range(3)]
    cfg = R.default_serve_config()
    cfg.checksum = 1
    srv = R.GpuServer(ctx, cfg)
    try:
        srv.pause()
        for i, f in enumerate(frames):
            srv.submit(f, i)
        R.inject_failure(R.FAIL_LAUNCH, 1)
        srv.pause(False)
        srv.drain()
        res = sorted(srv.poll(16), key=lambda r: r.tag)
        R.inject_failure(R.FAIL_LAUNCH, 0)
        assert [r.tag for r in res] == [0, 1, 2]
        assert all(r.status == R.RAV2D_EDEVICE for r in res if r.batch_frames == res[0].batch_frames)
        assert res[0].status == R.RAV2D_EDEVICE
        for i, f in enumerate(frames):
            srv.submit(f, 10 + i)
        srv.drain()
        res = sorted(srv.poll(16), key=lambda r: r.tag)
        assert [r.status for r in res] == [0, 0, 0]
        assert [(r.checksum[0], r.checksum[1]) for r in res] == [expect[i] for i in range(3)]
    finally:
        R.inject_failure(R.FAIL_LAUNCH, 0)
        srv.close()
        for f in frames:
            f.close()


@pytest.mark.gpu
def test_cancel_queued_jobs(ctx, sources):
    # jobs cancelled while queued come back with ECANCELED; the others finish
    # with the planes their frames give alone; a launched or unknown job
    # cannot be cancelled
    sfs, expect = sources
    frames = [R.GpuFrame.from_synth(ctx, sfs[i]) for i in range(5)]
    cfg = R.default_serve_config()
    cfg.checksum = 1
    srv = R.GpuServer(ctx, cfg)
    try:
        srv.pause()
        for i, f in enumerate(frames):
            srv.submit(f, 100 + i)
        assert srv.cancel(101) and srv.cancel(103)
        assert not srv.cancel(101) and not srv.cancel(999)
        srv.pause(False)
        srv.drain()
        res = {r.tag: r for r in srv.poll(16)}
        assert sorted(res) == [100, 101, 102, 103, 104]
        for i in range(5):
            r = res[100 + i]
            if i in (1, 3):
                assert r.status == R.RAV2D_ECANCELED
            else:
                assert r.status == 0 and (r.checksum[0], r.checksum[1]) == expect[i], i
        assert not srv.cancel(100)  # done
    finally:
        srv.close()
        for f in frames:
            f.close()


@pytest.mark.gpu
def test_injected_launch_failure_of_a_frame_leaves_it_usable(ctx, sources):
    sfs, expect = sources
    f = R.GpuFrame.from_synth(ctx, sfs[0])
    try:
        R.inject_failure(R.FAIL_LAUNCH, 1)
        with pytest.raises(R.Rav2dError, match="injected"):
            f.filter()
        f.filter()
        f.apply_grain()
        assert (f.checksum(1), f.checksum(2)) == expect[0]
    finally:
        R.inject_failure(R.FAIL_LAUNCH, 0)
        f.close()
