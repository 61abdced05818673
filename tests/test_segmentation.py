"""Segmented frames: per-segment deblock thresholds (init_deblock_thr_lut_y/uv,
crates/rav2d/src/deblock.rs:13-57, read through the segment maps in
setup_thr_cols/rows, :1477-1567) and lossless segments in CDEF and CCSO
(cdef.rs:894-926, 984-1025, 1117-1150; decode.rs:3868-3869).

The CPU tests pin the oracle's segmentation path by what the mathematics
fixes: a map of zero deltas is the unsegmented frame; one segment everywhere
with delta d is the unsegmented frame with every 64x64 qidx moved by d
(clipped as the per-segment LUT clips); an all-lossless segmented frame is left
alone by CDEF and CCSO, while the same masks without segmentation are not.
The GPU tests compare the HIP path with the oracle through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O

QIDX_OFF = 2 * 5120  # Av2Filter.qidx follows filter_y and filter_uv


def _params_copy(p):
    q = R.FrameParams()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(q))
    return q


def _oracle(sf, params, inloop, mask=None, seg=None):
    """Oracle chain with explicit params, optional replacement masks / maps."""
    lib = O.load()
    planes = sf.padded_planes()  # with the picture padding the deblocking reads
    bps = planes[0].itemsize
    mask_ptr = C.cast(mask, C.c_void_p) if mask is not None else sf.mask_ptr
    s, st, suv, uvst = seg if seg is not None else sf.segmaps()
    rc = lib.oracle_filter_frame_seg(C.byref(params), mask_ptr, sf.n_mask, sf.lr_ptr, sf.n_lr,
                                     planes[0].ctypes.data, planes[1].ctypes.data, planes[2].ctypes.data,
                                     planes[0].shape[1] * bps, planes[1].shape[1] * bps, inloop,
                                     s, st, suv, uvst)
    assert rc == 0
    return [a[:sf.plane_shape(pl)[0], :sf.plane_shape(pl)[1]] for pl, a in enumerate(planes)]


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("bd", [8, 10])
def test_zero_deltas_equal_unsegmented(bd):
    sf = R.SynthFrame(264, 200, bd, 0, seed=31, tools=R.TOOL_DEBLOCK, mask_flags=R.MASK_SEG)
    try:
        p = _params_copy(sf.params)
        for i in range(8):
            p.seg_delta_q[i] = 0
            p.seg_lossless[i] = 0
        ref = _oracle(sf, p, R.INLOOP_ALL)
        p.seg_enabled = 0
        plain = _oracle(sf, p, R.INLOOP_ALL)
        assert _same(ref, plain)
        assert not _same(ref, sf.planes())
    finally:
        sf.close()


@pytest.mark.parametrize("bd,delta", [(8, -40), (10, 37), (10, -200), (8, 120)])
def test_one_segment_equals_moved_qidx(bd, delta):
    sf = R.SynthFrame(264, 200, bd, 1, seed=32, tools=R.TOOL_DEBLOCK, mask_flags=R.MASK_SEG)
    try:
        p = _params_copy(sf.params)
        for i in range(8):
            p.seg_delta_q[i] = delta
            p.seg_lossless[i] = 0
        y, yst, uv, uvst = sf.segmaps()
        bw4, bh4 = yst, ((sf.height + 7) >> 3) * 2
        ymap = (C.c_uint8 * (yst * bh4))(*([5] * (yst * bh4)))
        uvmap = (C.c_uint8 * (uvst * (bh4 >> 1)))(*([5] * (uvst * (bh4 >> 1))))
        seg = (C.cast(ymap, C.c_void_p), yst, C.cast(uvmap, C.c_void_p), uvst)
        got = _oracle(sf, p, R.INLOOP_ALL, seg=seg)
        # the same frame unsegmented, every 64x64 qidx moved by delta
        qmax = 255 + 48 * (1 if bd > 8 else 0)
        raw = bytearray(bytes(sf.mask_bytes()))
        n = sf.n_mask
        for k in range(n):
            base = k * R.AV2FILTER_SIZE + QIDX_OFF
            q = np.frombuffer(bytes(raw[base:base + 32]), dtype=np.uint16).astype(np.int32)
            raw[base:base + 32] = np.clip(q + delta, 0, qmax).astype(np.uint16).tobytes()
        mask = (C.c_uint8 * len(raw)).from_buffer(raw)
        p.seg_enabled = 0
        want = _oracle(sf, p, R.INLOOP_ALL, mask=mask)
        assert _same(got, want)
    finally:
        sf.close()


@pytest.mark.parametrize("bd,sb128", [(8, 0), (10, 1)])
def test_segments_per_64x64_equal_moved_qidx_per_64x64(bd, sb128):
    """A segment per 64x64 block with its own delta is the unsegmented frame
    with each 64x64 qidx moved by its segment's delta: pins where the maps are
    read, across 64x64 and superblock boundaries and for chroma."""
    w, h = 328, 264
    sf = R.SynthFrame(w, h, bd, sb128, seed=34, tools=R.TOOL_DEBLOCK, mask_flags=R.MASK_SEG)
    try:
        rng = np.random.default_rng(5)
        p = _params_copy(sf.params)
        deltas = rng.integers(-90, 90, 8)
        for i in range(8):
            p.seg_delta_q[i] = int(deltas[i])
            p.seg_lossless[i] = 0
        _, yst, _, uvst = sf.segmaps()
        bh4 = ((h + 7) >> 3) * 2
        n64x, n64y = (yst + 15) // 16, (bh4 + 15) // 16
        seg64 = rng.integers(0, 8, (n64y, n64x))
        ymap = np.repeat(np.repeat(seg64, 16, 0), 16, 1)[:bh4, :yst].astype(np.uint8).copy()
        uvmap = np.repeat(np.repeat(seg64, 8, 0), 8, 1)[:bh4 >> 1, :uvst].astype(np.uint8).copy()
        seg = (ymap.ctypes.data, yst, uvmap.ctypes.data, uvst)
        got = _oracle(sf, p, R.INLOOP_ALL, seg=seg)
        qmax = 255 + 48 * (1 if bd > 8 else 0)
        raw = bytearray(bytes(sf.mask_bytes()))
        sbw = (yst + 63) // 64
        for k in range(sf.n_mask):
            base = k * R.AV2FILTER_SIZE + QIDX_OFF
            q = np.frombuffer(bytes(raw[base:base + 32]), dtype=np.uint16).astype(np.int32).copy()
            for j in range(16):
                y64, x64 = (k // sbw) * 4 + j // 4, (k % sbw) * 4 + j % 4
                if y64 < n64y and x64 < n64x:
                    q[j] = np.clip(q[j] + deltas[seg64[y64, x64]], 0, qmax)
            raw[base:base + 32] = q.astype(np.uint16).tobytes()
        mask = (C.c_uint8 * len(raw)).from_buffer(raw)
        p.seg_enabled = 0
        want = _oracle(sf, p, R.INLOOP_ALL, mask=mask)
        assert _same(got, want)
        assert not _same(got, _oracle(sf, p, R.INLOOP_ALL))
    finally:
        sf.close()


def test_lossless_segment_skips_cdef_and_ccso():
    tools = R.TOOL_CDEF | R.TOOL_CCSO
    sf = R.SynthFrame(200, 136, 10, 0, seed=33, tools=tools, lossless_pct=100, skip_pct=0,
                      mask_flags=R.MASK_SEG)
    try:
        p = _params_copy(sf.params)
        assert p.seg_lossless[7] == 1
        inp = sf.planes()
        assert _same(_oracle(sf, p, R.INLOOP_ALL), inp)
        # the same lossless masks on an unsegmented frame are not read by CDEF
        p.seg_enabled = 0
        assert not _same(_oracle(sf, p, R.INLOOP_ALL), inp)
    finally:
        sf.close()


# ---- GPU vs oracle ----------------------------------------------------------

CASES = [
    # (w, h, bitdepth, sb128, seed, lossless_pct, tools, mask_flags)
    (352, 288, 8, 0, 41, 10, R.TOOL_DEBLOCK, R.MASK_SEG),
    (640, 360, 10, 1, 42, 15, R.TOOL_DEBLOCK, R.MASK_SEG | R.MASK_TXPART | R.MASK_SUBPU),
    (352, 288, 10, 0, 43, 20, R.TOOL_CDEF | R.TOOL_CCSO, R.MASK_SEG),
    (264, 136, 8, 1, 44, 30, R.TOOL_CDEF | R.TOOL_CCSO, R.MASK_SEG),
    (640, 360, 10, 0, 45, 10, R.TOOL_ALL & ~R.TOOL_GRAIN, R.MASK_SEG | R.MASK_SUBPU),
    (1920, 1080, 10, 1, 46, 5, R.TOOL_ALL & ~R.TOOL_GRAIN, R.MASK_SEG | R.MASK_TXPART | R.MASK_SUBPU),
]


@pytest.fixture(scope="module")
def ctx():
    c = R.GpuContext(0)
    yield c
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d_%db_sb%d_s%d_ll%d_t%d_m%d" % (
    c[0], c[1], c[2], 64 << c[3], c[4], c[5], c[6], c[7]))
def test_segmented_frame_matches_oracle(ctx, case):
    w, h, bd, sb128, seed, ll, tools, flags = case
    sf = R.SynthFrame(w, h, bd, sb128, seed, tools=tools, lossless_pct=ll, mask_flags=flags)
    f = R.GpuFrame.from_synth(ctx, sf)
    try:
        ref = O.filter_frame(sf)
        f.filter()
        got = f.download(1)
        for pl in range(3):
            if not np.array_equal(ref[pl], got[pl]):
                d = np.argwhere(ref[pl] != got[pl])
                y, x = d[0]
                raise AssertionError("plane %d: %d mismatches, first at (y=%d, x=%d): ref %d got %d"
                                     % (pl, len(d), y, x, ref[pl][y, x], got[pl][y, x]))
        assert not _same(got, sf.planes())
    finally:
        f.close()
        sf.close()


@pytest.mark.gpu
def test_segment_map_required_and_checked(ctx):
    sf = R.SynthFrame(136, 72, 10, 0, seed=47, tools=R.TOOL_DEBLOCK, mask_flags=R.MASK_SEG)
    f = R.GpuFrame(ctx, sf.params)
    try:
        f.set_side_info(sf.mask_ptr, sf.n_mask, sf.lr_ptr, sf.n_lr)
        with pytest.raises(R.Rav2dError) as e:
            f.filter()
        assert "segment map" in str(e.value)
        y, yst, uv, uvst = sf.segmaps()
        bad = (C.c_uint8 * (yst * 18))(*([8] * (yst * 18)))
        with pytest.raises(R.Rav2dError):
            f.set_segmap(C.cast(bad, C.c_void_p), yst, uv, uvst)
        with pytest.raises(R.Rav2dError):
            f.set_segmap(y, yst - 1, uv, uvst)
        f.set_segmap(y, yst, uv, uvst)
        f.filter()
    finally:
        f.close()
        sf.close()
