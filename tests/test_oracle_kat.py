"""The oracle against the reference's own known answers and against
invariants its unit tests state (CPU only)."""
import ctypes as C

import numpy as np

import rav2d_amd as R
import oracle_lib as O


def test_dq_lookup_closed_form():
    lib = O.load()
    # quantizer.rs:31-47: 64 at 0, table value << (q-1)/24 otherwise
    assert lib.oracle_dq_lookup(0) == 64
    assert lib.oracle_dq_lookup(1) == 40
    assert lib.oracle_dq_lookup(24) == 78
    assert lib.oracle_dq_lookup(25) == 80          # 40 << 1
    tbl = [40, 41, 43, 44, 45, 47, 48, 49, 51, 52, 54, 55, 57, 59, 60, 62, 64, 66, 68, 70, 72, 74, 76, 78]
    for q in range(1, 256):
        assert lib.oracle_dq_lookup(q) == tbl[(q - 1) % 24] << ((q - 1) // 24)
    assert lib.oracle_dq_lookup(-5) == 64 and lib.oracle_dq_lookup(999) == lib.oracle_dq_lookup(255)


def test_deblock_quant_thr_kat():
    lib = O.load()
    # lf_mask.rs:830: deblock_quant_thr(0, 0) == (64 + 4) >> 9
    assert lib.oracle_deblock_quant_thr(0, 0) == (64 + 4) >> 9
    # closed form (lf_mask.rs:332-335) incl. the qmax clamp for hbd
    for hbd in (0, 1):
        for q in range(-10, 320):
            qq = min(max(q, 0), 255 + 48 * hbd)
            qq = min(max(qq, 0), 255)
            dq = 64 if qq == 0 else [40, 41, 43, 44, 45, 47, 48, 49, 51, 52, 54, 55, 57, 59, 60, 62, 64,
                                     66, 68, 70, 72, 74, 76, 78][(qq - 1) % 24] << ((qq - 1) // 24)
            assert lib.oracle_deblock_quant_thr(hbd, q) == (dq + 4) >> 9


def test_deblock_side_thr_table_endpoints():
    lib = O.load()
    # DEBLOCK_SIDE_THRESHOLDS[0] = -16 -> max(-16 + 16, 0) >> 5 = 0 at 8 bit
    assert lib.oracle_deblock_side_thr(0, 0) == 0
    # last entry 1678: (1678 + 16) >> 5 = 52 ; 10 bit: q_ind = q - 48, (1678 + (16 >> 2)) >> 3
    assert lib.oracle_deblock_side_thr(0, 295) == (1678 + 16) >> 5
    assert lib.oracle_deblock_side_thr(1, 295 + 48) == (1678 + 4) >> 3
    # monotone non-decreasing in qidx
    v = [lib.oracle_deblock_side_thr(0, q) for q in range(296)]
    assert all(a <= b for a, b in zip(v, v[1:]))


def test_uniform_frame_unchanged_by_deblock():
    """deblock.rs:2409-2427: a uniform image is left unchanged."""
    sf = R.SynthFrame(128, 128, 10, 0, seed=9, tools=R.TOOL_DEBLOCK)
    for pl in range(3):
        h, w = sf.plane_shape(pl)
        arr = np.frombuffer((C.c_uint16 * (h * w)).from_address(sf.plane_ptr(pl)), dtype=np.uint16)
        arr[:] = 512
    out = O.filter_frame(sf, R.INLOOP_DEBLOCK)
    for pl in range(3):
        assert (out[pl] == 512).all()


def test_deblock_all_lossless_unchanged():
    """deblock.rs:2463-2490: lossless sides are never written."""
    sf = R.SynthFrame(256, 128, 8, 0, seed=10, tools=R.TOOL_DEBLOCK, lossless_pct=100)
    inp = sf.planes()
    out = O.filter_frame(sf, R.INLOOP_DEBLOCK)
    for a, b in zip(inp, out):
        assert np.array_equal(a, b)


def test_deblock_changes_blocky_frame():
    sf = R.SynthFrame(256, 256, 8, 0, seed=11, tools=R.TOOL_DEBLOCK)
    inp = sf.planes()
    out = O.filter_frame(sf, R.INLOOP_DEBLOCK)
    assert (inp[0] != out[0]).sum() > 1000
    # deblocking only moves samples next to transform edges: interior of 4x4 grid untouched?
    # (edges are at multiples of 4; the filter reaches at most 8 samples across)
    assert np.abs(inp[0].astype(int) - out[0]).max() < 256
