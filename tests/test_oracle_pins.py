"""Pins of the oracle's filter arithmetic that do not restate the oracle.

Each check is written independently of the oracle's code and chosen so that a
plausible mistake in the restatement (a wrong tap offset, sign, rounding,
ring-buffer row or edge rule) fails it:

* the multi-class / PC-Wiener class assignment (get_class_lut_idx,
  looprestoration.rs:194-271, and get_qval_given_tskip, :40-49) restated in
  numpy, each class given its own random filter so the output shows which
  filter every 4x4 block took;
* loop restoration FIRs (NS-Wiener luma, NS-Wiener chroma with its
  cross-component luma taps, NS multi-class and PC-Wiener filtering): the
  reference's own scalar closed forms (simd.rs:1129-1224:
  s = m*128 + sum (a + b - 2m) c, PC: s = m*c12 + sum (a + b) c,
  out = clip((s + 64) >> 7)) applied in numpy to a "virtual image" built here
  from the reference's row/column rules (backup_row / backup_row_lpf /
  backup_row_luma, looprestoration.rs:54-109, 408-522, and the stripe row
  walk of ns_wiener_single_y / ns_wiener_single_uv, :524-713, :1145-1430),
  with the tap tables read from the reference as data
  (tests/golden/lr_taps.json);
* CDEF block filter: the AV1 specification's CDEF filter process (tap
  weights {4,2}/{3,3} and {2,1}, constrain with damping adjustment, rounding
  x + ((8 + sum - (sum < 0)) >> 4), clipping to the taps' range) on interior
  blocks, where rav2d's filter (cdef.rs:320-460) is the same computation;
* deblock_bd (deblock.rs:217-283): mirror symmetry across the edge (up to
  its round-half-up ties), DC-shift equivariance, invariance on linear ramps,
  corrections tapering away from the edge and the reference's own
  known-answer tests (deblock.rs:2386-2490);
* GDF error (gdf_prep, looprestoration.rs:1432-1643): a numpy restatement of
  the reference's computation (10-bit domain, 2x2 gradient classes over the
  4x4 gradient window, the 18 clipped symmetric taps and 4 gradient features
  weighted into three quantised indices, the error-table lookup) on every
  reference / qp pair and bit depth, on random, low-contrast and diagonal
  textures (the last two make the class comparisons tie); plus invariance to a
  DC shift and the 8-bit path equal to the 10-bit path on samples x 4;
* every lookup table in rav2d_amd/csrc/tables.h against digests of the
  reference's Rust array literals taken by a parser of its own
  (tests/golden/table_digests.py)."""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as O
import tables_h

HERE = os.path.dirname(os.path.abspath(__file__))
TAPS = json.load(open(os.path.join(HERE, "golden", "lr_taps.json")))
TOP, BOTTOM, LEFT, RIGHT = 4, 8, 1, 2


def _lib():
    lib = O.load()
    P = C.c_void_p
    if not getattr(lib, "_pins", False):
        lib.oracle_lr_ns_wiener_y.argtypes = [P, C.c_ssize_t, P, P, P, C.c_int, C.c_int, P, C.c_int, P, C.c_int]
        lib.oracle_lr_ns_wiener_y.restype = None
        lib.oracle_lr_wiener_multi.argtypes = [P, C.c_ssize_t, P, P, P, C.c_int, C.c_int, P, P, P, P, C.c_int,
                                               C.c_int, P, C.c_int]
        lib.oracle_lr_wiener_multi.restype = None
        lib.oracle_lr_ns_wiener_uv.argtypes = [P, C.c_ssize_t, P, P, P, C.c_int, C.c_int, P, P, C.c_ssize_t, P, P,
                                               C.c_int, C.c_int, C.c_int, C.c_int, P, C.c_int]
        lib.oracle_lr_ns_wiener_uv.restype = None
        lib.oracle_lr_gdf_prep.argtypes = [P, C.c_int, P, C.c_ssize_t, P, P, P, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_int]
        lib.oracle_lr_gdf_prep.restype = None
        lib.oracle_lr_taps.argtypes = [C.c_int, P]
        lib.oracle_lr_taps.restype = C.c_int
        lib.oracle_cdef_filter_block_16.argtypes = [P, C.c_ssize_t, P, P, P, C.c_int, C.c_int, C.c_int, C.c_int,
                                                    C.c_int, C.c_int, C.c_int, C.c_int]
        lib.oracle_cdef_filter_block_16.restype = None
        lib.oracle_deblock_edge_16.argtypes = [P, C.c_ssize_t, C.c_int, C.c_uint32, C.c_uint32, C.c_ssize_t,
                                               C.c_ssize_t, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.oracle_deblock_edge_16.restype = None
        lib._pins = True
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def test_tap_tables_are_the_references():
    lib = _lib()
    names = ["WIENER_NS_CONFIG_Y", "WIENER_NS_CONFIG_UV", "WIENER_NS_CONFIG_UV_FROM_Y", "PC_WIENER_CONFIG",
             "GDF_COORDS"]
    for which, name in enumerate(names):
        buf = np.zeros(64, dtype=np.int8)
        n = lib.oracle_lr_taps(which, _p(buf))
        assert buf[:2 * n].reshape(n, 2).tolist() == TAPS[name], name


def test_lookup_tables_are_the_references():
    digests = json.load(open(os.path.join(HERE, "golden", "table_digests.json")))
    tabs = tables_h.tables()
    assert sorted(tabs) == sorted(digests)
    for name, want in digests.items():
        vals = tabs[name]
        assert len(vals) == want["n"], name
        assert hashlib.sha256(vals.astype("<i8").tobytes()).hexdigest() == want["sha256"], name


# ---------------------------------------------------------------- LR stripes
class Stripe:
    """One stripe of one restoration unit in the layout the reference's
    stripe functions take: the stripe rows P (w samples plus 8 beyond each
    side), the 6 samples left of each row (a separate buffer, as the
    reference's left backups), two rows above (lpf) and below (lpf_bottom).
    Every buffer holds independent random samples, so reading the wrong one
    shows."""

    PAD = 8

    def __init__(self, rng, w, h, bitdepth, edges, hi=None):
        hi = hi or 1 << bitdepth
        self.w, self.h, self.bd, self.edges = w, h, bitdepth, edges
        self.stride = w + 2 * self.PAD
        self.P = rng.integers(0, hi, size=(h, self.stride), dtype=np.uint16)
        self.lpf = rng.integers(0, hi, size=(2, self.stride), dtype=np.uint16)
        self.lpfb = rng.integers(0, hi, size=(2, self.stride), dtype=np.uint16)
        self.left = rng.integers(0, hi, size=(h, 6), dtype=np.uint16)

    def ptrs(self):
        return (self.P.ctypes.data + 2 * self.PAD, self.lpf.ctypes.data + 2 * self.PAD,
                self.lpfb.ctypes.data + 2 * self.PAD)

    def row(self, r, edge_len):
        """Row r of the virtual image over columns [-edge_len, w + edge_len),
        following backup_row (stripe rows: the left buffer beyond the left
        edge) and backup_row_lpf (rows above / below: their own samples);
        without an edge flag the first / last sample repeats."""
        w, pad, e = self.w, self.PAD, self.edges
        if 0 <= r < self.h:
            src, left = self.P[r], self.left[r]
        else:
            src, left = self.outside_row(r), None
        mid = src[pad:pad + w].astype(np.int64)
        if e & LEFT:
            lft = left[6 - edge_len:].astype(np.int64) if left is not None else \
                src[pad - edge_len:pad].astype(np.int64)
        else:
            lft = np.full(edge_len, mid[0])
        rgt = src[pad + w:pad + w + edge_len].astype(np.int64) if e & RIGHT else np.full(edge_len, mid[-1])
        return np.concatenate([lft, mid, rgt])

    def outside_row(self, r):
        e = self.edges
        if r < 0:
            if e & TOP:
                return self.lpf[1] if r == -1 else self.lpf[0]
            return self.P[0]
        if e & BOTTOM:
            return self.lpfb[0] if r == self.h else self.lpfb[1]
        return self.P[self.h - 1]

    def image(self, reach, edge_len):
        """Rows [-reach, h + reach) of the virtual image, columns
        [-edge_len, w + edge_len)."""
        rows = []
        for r in range(-reach, self.h + reach):
            if r < 0 and not (self.edges & TOP):
                rr = 0
            elif r >= self.h and not (self.edges & BOTTOM):
                rr = self.h - 1
            else:
                rr = r
            rows.append(self.row(rr, edge_len))
        return np.array(rows)


def _ll(rng, h, w, frac=0.2):
    ll = np.zeros((max(h // 4, 1) + 1, 4), dtype=np.uint16)
    for r in range(h // 4):
        bits = 0
        for b in range(w // 4):
            if rng.random() < frac:
                bits |= 1 << b
        ll[r, 0] = bits
    return ll


def _apply_ll(out, orig, ll, w, h):
    for by in range(h // 4):
        for bx in range(w // 4):
            if ll[by, 0] >> bx & 1:
                out[by * 4:by * 4 + 4, bx * 4:bx * 4 + 4] = orig[by * 4:by * 4 + 4, bx * 4:bx * 4 + 4]
    return out


def _ns_closed_form(V, reach, edge_len, taps, coefs, w, h, bd):
    """s = m * 128 + sum_i (V[y+dy][x+dx] + V[y-dy][x-dx] - 2m) c_i, clip((s + 64) >> 7)."""
    m = V[reach:reach + h, edge_len:edge_len + w]
    s = m * 128
    for (dy, dx), c in zip(taps, coefs):
        a = V[reach + dy:reach + dy + h, edge_len + dx:edge_len + dx + w]
        b = V[reach - dy:reach - dy + h, edge_len - dx:edge_len - dx + w]
        s = s + (a + b - 2 * m) * int(c)
    return np.clip((s + 64) >> 7, 0, (1 << bd) - 1)


def _pc_closed_form(V, reach, edge_len, taps, coefs, center, w, h, bd):
    """s = m * c12 + sum_i (V[y+dy][x+dx] + V[y-dy][x-dx]) c_i, clip((s + 64) >> 7)."""
    m = V[reach:reach + h, edge_len:edge_len + w]
    s = m * int(center)
    for (dy, dx), c in zip(taps, coefs):
        a = V[reach + dy:reach + dy + h, edge_len + dx:edge_len + dx + w]
        b = V[reach - dy:reach - dy + h, edge_len - dx:edge_len - dx + w]
        s = s + (a + b) * int(c)
    return np.clip((s + 64) >> 7, 0, (1 << bd) - 1)


LR_GEOMS = [(64, 64), (64, 56), (40, 16), (20, 8), (64, 60)]


@pytest.mark.parametrize("edges", range(16))
@pytest.mark.parametrize("bd", [8, 10, 12])
def test_ns_wiener_luma_is_the_closed_form(edges, bd):
    lib = _lib()
    rng = np.random.default_rng(1000 + edges * 7 + bd)
    taps = TAPS["WIENER_NS_CONFIG_Y"]
    for w, h in LR_GEOMS:
        st = Stripe(rng, w, h, bd, edges)
        filt = rng.integers(-128, 128, size=16).astype(np.int8)
        ll = _ll(rng, h, w)
        V = st.image(4, 4)
        orig = st.P[:, 8:8 + w].astype(np.int64).copy()
        want = _apply_ll(_ns_closed_form(V, 4, 4, taps, filt, w, h, bd), orig, ll, w, h)
        p, lpf, lpfb = st.ptrs()
        lib.oracle_lr_ns_wiener_y(p, st.stride, _p(st.left), lpf, lpfb, w, h, _p(filt), edges, _p(ll), bd)
        got = st.P[:, 8:8 + w].astype(np.int64)
        assert np.array_equal(got, want), (w, h, edges, bd)


@pytest.mark.parametrize("edges", [0, 15, TOP | LEFT, BOTTOM | RIGHT, TOP | BOTTOM, LEFT | RIGHT])
@pytest.mark.parametrize("bd", [8, 10, 12])
def test_ns_multi_class_filtering_is_the_closed_form(edges, bd):
    # every class carries the same filter, so the output does not depend on
    # the class assignment: this pins the multi-class path's FIR and rows
    lib = _lib()
    rng = np.random.default_rng(2000 + edges * 5 + bd)
    taps = TAPS["WIENER_NS_CONFIG_Y"]
    for w, h in LR_GEOMS:
        st = Stripe(rng, w, h, bd, edges)
        filt = rng.integers(-128, 128, size=16).astype(np.int8)
        user = np.zeros((16, 18), dtype=np.int8)
        user[:, :16] = filt
        sub = rng.integers(0, 16, size=256).astype(np.uint8)
        noskip = rng.integers(0, 1 << 16, size=20).astype(np.uint16)
        ll = _ll(rng, h, w)
        V = st.image(4, 4)
        orig = st.P[:, 8:8 + w].astype(np.int64).copy()
        want = _apply_ll(_ns_closed_form(V, 4, 4, taps, filt, w, h, bd), orig, ll, w, h)
        p, lpf, lpfb = st.ptrs()
        lib.oracle_lr_wiener_multi(p, st.stride, _p(st.left), lpf, lpfb, w, h, _p(user), None, _p(sub),
                                   _p(noskip), int(rng.integers(4, 1500)), edges, _p(ll), bd)
        assert np.array_equal(st.P[:, 8:8 + w].astype(np.int64), want), (w, h, edges, bd)


def _qval_given_tskip(qstep, tskip, i, bdm8):
    """get_qval_given_tskip (looprestoration.rs:40-49) in numpy: qstep is
    brought to 8-bit precision with rounding, the mode's three weights
    (MODE_WEIGHTS[i]) apply to tskip << 5, qstep and (tskip qstep + 128) >> 8,
    the sum is rounded by 2^13 away from zero and offset by MODE_OFFSETS[i];
    the result is 255 times that."""
    mw = np.array(TAPS["MODE_WEIGHTS"], dtype=np.int64).reshape(4, 3)[i]
    qstep = (np.asarray(qstep, dtype=np.int64) + ((1 << bdm8) >> 1)) >> bdm8
    tskip = np.asarray(tskip, dtype=np.int64)
    prod = (tskip * qstep + 128) >> 8
    q = mw[0] * (tskip << 5) + mw[1] * qstep + mw[2] * prod
    q = np.sign(q) * ((np.abs(q) + (1 << 12)) >> 13)
    return 255 * (TAPS["MODE_OFFSETS"][i] + q)


def test_get_qval_given_tskip_is_the_closed_form():
    L = O.load()
    qs, ts = np.meshgrid(np.arange(0, 1601), np.arange(0, 253, 3), indexing="ij")
    for bdm8 in (0, 2, 4):
        for i in range(4):
            want = _qval_given_tskip(qs, ts, i, bdm8)
            got = np.array([[L.oracle_get_qval_given_tskip(int(q), int(t), i, bdm8) for t in ts[0]]
                            for q in qs[:, 0]])
            assert np.array_equal(got, want), (i, bdm8)


def _class_lut_index(st, noskip, base_q):
    """The class-LUT index of every 4x4 block of a stripe (get_class_lut_idx,
    looprestoration.rs:194-271, as wiener_multi calls it, :840-895).  Over the
    6x6 samples from (-1, -1) to (4, 4) of the block, sums of |second
    differences| vertical, anti-diagonal (up-right / down-left) and diagonal
    (up-left / down-right), each times PC_WIENER_NORMALIZER[i], rounded back
    to 8 bits; the skip count weighs the 3x3 neighbouring blocks of the
    noskip mask (centre 16, edge 4, corner 1, indexes clamped to the mask)
    times PC_WIENER_NORMALIZER[3].  Four 3-bit digits
    min((max(0, x) + 2^13) >> 14, 255) >> 5 with x = qval(mode 0) and
    x = f_i + qval(mode i + 1) form the index.  Rows: the stripe's virtual
    image, except that the last block row reads its two rows below the block
    as the rows below the stripe (lpf_bottom) or, without them, as the
    block's last row, whatever the stripe's height."""
    w, h, bd = st.w, st.h, st.bd
    bh, bw = h >> 2, w >> 2
    bdm8 = bd - 8
    norm = TAPS["PC_WIENER_NORMALIZER"]
    V = st.image(4, 4).astype(np.int64)
    out = np.zeros((bh, bw), dtype=np.int64)
    for by in range(bh):
        rows = [V[4 + r] for r in range(4 * by - 2, 4 * by + 6)]
        if by + 1 == bh:
            if st.edges & BOTTOM:
                rows[6], rows[7] = st.row(h, 4), st.row(h + 1, 4)
            else:
                rows[6] = rows[7] = rows[5]
        R = np.array(rows)  # row 4by - 2 + k; column c - 4
        c0, n = 3, 4 * bw + 2  # columns -1 .. 4bw
        m, up, dn = R[1:7, c0:c0 + n], R[0:6], R[2:8]
        feats = [np.abs(up[:, c0:c0 + n] - 2 * m + dn[:, c0:c0 + n]).sum(0),
                 np.abs(up[:, c0 + 1:c0 + 1 + n] - 2 * m + dn[:, c0 - 1:c0 - 1 + n]).sum(0),
                 np.abs(up[:, c0 - 1:c0 - 1 + n] - 2 * m + dn[:, c0 + 1:c0 + 1 + n]).sum(0)]
        for bx in range(bw):
            f = [(int(a[4 * bx:4 * bx + 6].sum()) * norm[i] + ((1 << bdm8) >> 1)) >> bdm8
                 for i, a in enumerate(feats)]
            skipped = 0
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    fx = min(max((bx & 15) + dx, 0), 15)
                    fy = min(max(by + dy, 0), bh - 1)
                    if not (int(noskip[fy]) >> fx) & 1:
                        skipped += (16, 4, 1)[(dy != 0) + (dx != 0)]
            tskip = skipped * norm[3]

            def digit(x):
                return min((max(0, int(x)) + (1 << 13)) >> 14, 255) >> 5
            idx = digit(_qval_given_tskip(base_q, tskip, 0, bdm8)) << 9
            for i in range(3):
                idx |= digit(f[i] + _qval_given_tskip(base_q, tskip, i + 1, bdm8)) << (3 * (2 - i))
            out[by, bx] = idx
    return out


def _multi_closed_form(st, filters, pc, subclass, noskip, base_q, ll):
    """wiener_multi's output: each 4x4 block not in the skip mask filtered
    with filters[subclass[PC_WIENER_LUT_TO_CLASS[index]]] (the NS or the
    PC-Wiener closed form); rows and columns beyond the last whole block are
    left as they are."""
    cls = tables_h.tables()["R2T_PC_WIENER_LUT_TO_CLASS"][_class_lut_index(st, noskip, base_q)]
    which = subclass[cls]
    w, h, bd = st.w, st.h, st.bd
    V = st.image(4, 4).astype(np.int64)
    out = st.P[:, 8:8 + w].astype(np.int64).copy()
    taps = TAPS["PC_WIENER_CONFIG"] if pc else TAPS["WIENER_NS_CONFIG_Y"]
    for by in range(h >> 2):
        for bx in range(w >> 2):
            if ll[by, 0] >> bx & 1:
                continue
            F = filters[which[by, bx]]
            sub = V[4 * by:4 * by + 12, 4 * bx:4 * bx + 12]
            blk = _pc_closed_form(sub, 4, 4, taps, F[:12], F[12], 4, 4, bd) if pc else \
                _ns_closed_form(sub, 4, 4, taps, F[:16], 4, 4, bd)
            out[4 * by:4 * by + 4, 4 * bx:4 * bx + 4] = blk
    return out


@pytest.mark.parametrize("pc", [0, 1], ids=["ns_multi", "pc_wiener"])
@pytest.mark.parametrize("edges", [0, 15, TOP | LEFT, BOTTOM | RIGHT, TOP | BOTTOM])
@pytest.mark.parametrize("bd", [8, 10, 12])
def test_multi_class_selection_is_the_closed_form(bd, edges, pc):
    """Every class carries its own random filter, so the output shows which
    filter each block took.  Low-contrast and diagonal textures keep the
    features inside the quantiser's range (full-range noise saturates every
    digit); stripe heights 62 and 30 exercise the last block row's rows."""
    lib = _lib()
    rng = np.random.default_rng(6000 + bd * 13 + edges * 2 + pc)
    for w, h, kind in [(64, 64, "random"), (64, 64, "low"), (64, 56, "low"), (64, 62, "low"), (36, 30, "low"),
                       (40, 16, "diagonal"), (20, 8, "low")]:
        st = Stripe(rng, w, h, bd, edges, hi=(24 << (bd - 8)) if kind == "low" else None)
        if kind == "diagonal":
            _diagonal_texture(rng, st)
        if pc:
            filters = rng.integers(-400, 400, size=(64, 13)).astype(np.int16)
            subclass = (np.arange(256) & 63).astype(np.uint8)
        else:
            filters = rng.integers(-128, 128, size=(16, 18)).astype(np.int8)
            subclass = rng.integers(0, 16, size=256).astype(np.uint8)
        noskip = rng.integers(0, 1 << 16, size=20).astype(np.uint16)
        base_q = int(rng.integers(4, 1500))
        ll = _ll(rng, h, w)
        want = _multi_closed_form(st, filters, pc, subclass, noskip, base_q, ll)
        p, lpf, lpfb = st.ptrs()
        lib.oracle_lr_wiener_multi(p, st.stride, _p(st.left), lpf, lpfb, w, h, None if pc else _p(filters),
                                   _p(filters) if pc else None, _p(subclass), _p(noskip), base_q, edges, _p(ll), bd)
        assert np.array_equal(st.P[:, 8:8 + w].astype(np.int64), want), (w, h, kind)


@pytest.mark.parametrize("edges", [0, 15, TOP | LEFT, BOTTOM | RIGHT, TOP | BOTTOM, LEFT | RIGHT])
@pytest.mark.parametrize("bd", [8, 10, 12])
def test_pc_wiener_filtering_is_the_closed_form(edges, bd):
    lib = _lib()
    rng = np.random.default_rng(3000 + edges * 5 + bd)
    taps = TAPS["PC_WIENER_CONFIG"]
    for w, h in LR_GEOMS:
        st = Stripe(rng, w, h, bd, edges)
        coefs = rng.integers(-400, 400, size=12)
        center = int(rng.integers(-2000, 2000))
        pre = np.zeros((64, 13), dtype=np.int16)
        pre[:, :12] = coefs
        pre[:, 12] = center
        sub = rng.integers(0, 64, size=256).astype(np.uint8)
        noskip = rng.integers(0, 1 << 16, size=20).astype(np.uint16)
        ll = _ll(rng, h, w)
        V = st.image(4, 4)
        orig = st.P[:, 8:8 + w].astype(np.int64).copy()
        want = _apply_ll(_pc_closed_form(V, 4, 4, taps, coefs, center, w, h, bd), orig, ll, w, h)
        p, lpf, lpfb = st.ptrs()
        lib.oracle_lr_wiener_multi(p, st.stride, _p(st.left), lpf, lpfb, w, h, None, _p(pre), _p(sub),
                                   _p(noskip), int(rng.integers(4, 1500)), edges, _p(ll), bd)
        assert np.array_equal(st.P[:, 8:8 + w].astype(np.int64), want), (w, h, edges, bd)


class LumaOfStripe:
    """The luma the chroma stripe's cross-component taps read (4:2:0): rows
    2y, 2y + 1 of the stripe's luma for chroma row y; above the stripe the
    two deblocked luma rows (top), below it two more (bottom), each buffer
    independent and random, 8 luma samples beyond each side."""

    def __init__(self, rng, cw, ch, bd):
        hi = 1 << bd
        self.cw, self.ch = cw, ch
        self.stride = 2 * cw + 16
        self.L = rng.integers(0, hi, size=(2 * ch, self.stride), dtype=np.uint16)
        self.top = rng.integers(0, hi, size=(2, self.stride), dtype=np.uint16)
        self.bot = rng.integers(0, hi, size=(2, self.stride), dtype=np.uint16)

    def down(self, r0, r1, ds):
        """Downsampled luma row from luma rows r0, r1 at even luma columns
        (backup_row_luma): ds 0 the 2x2 average, 1 the vertical pair, 2 the
        top-left sample; index i = luma column i - 8."""
        a, b = r0.astype(np.int64), r1.astype(np.int64)
        if ds == 0:
            v = (a[:-1] + b[:-1] + a[1:] + b[1:]) >> 2
            return np.concatenate([v, [0]])
        if ds == 1:
            return (a + b) >> 1
        return a

    def image(self, edges, ds):
        """Rows [-2, ch + 2) of downsampled luma at chroma columns
        [-2, cw + 2) (luma column 2x)."""
        cw, ch = self.cw, self.ch
        out = []
        for r in range(-2, ch + 2):
            if r < 0:
                if edges & TOP:
                    row = self.down(self.top[0], self.top[0], ds) if r == -2 else \
                        self.down(self.top[0], self.top[1], ds)
                else:
                    row = self.down(self.L[0], self.L[1], ds)
            elif r >= ch:
                if edges & BOTTOM:
                    row = self.down(self.bot[0], self.bot[1], ds) if r == ch else \
                        self.down(self.bot[1], self.bot[1], ds)
                else:
                    row = self.down(self.L[2 * ch - 2], self.L[2 * ch - 1], ds)
            else:
                row = self.down(self.L[2 * r], self.L[2 * r + 1], ds)
            mid = row[8:8 + 2 * cw]
            if edges & LEFT:
                lft = row[8 - 4:8]
            else:
                lft = np.full(4, mid[0])
            if edges & RIGHT:
                rgt = row[8 + 2 * cw:8 + 2 * cw + 4]
            else:
                rgt = np.full(4, mid[2 * cw - 2])
            full = np.concatenate([lft, mid, rgt])  # luma columns -4 .. 2cw + 3
            out.append(full[0::2])  # luma columns -4, -2, .. -> chroma columns -2 .. cw + 1
        return np.array(out)


@pytest.mark.parametrize("edges", range(16))
@pytest.mark.parametrize("ds", [0, 1, 2])
def test_ns_wiener_chroma_is_the_closed_form(edges, ds):
    lib = _lib()
    bd = 10 if edges & 1 else 8
    rng = np.random.default_rng(4000 + edges * 3 + ds)
    taps_uv = TAPS["WIENER_NS_CONFIG_UV"]
    taps_y = TAPS["WIENER_NS_CONFIG_UV_FROM_Y"]
    for w, h in [(32, 32), (32, 28), (20, 8), (12, 16)]:
        st = Stripe(rng, w, h, bd, edges)
        lu = LumaOfStripe(rng, w, h, bd)
        filt = rng.integers(-128, 128, size=18).astype(np.int8)
        ll = _ll(rng, h, w)
        V = st.image(2, 2)
        LV = lu.image(edges, ds)
        m = V[2:2 + h, 2:2 + w]
        s = m * 128
        for (dy, dx), c in zip(taps_uv, filt[:6]):
            a = V[2 + dy:2 + dy + h, 2 + dx:2 + dx + w]
            b = V[2 - dy:2 - dy + h, 2 - dx:2 - dx + w]
            s = s + (a + b - 2 * m) * int(c)
        lc = LV[2:2 + h, 2:2 + w]
        for (dy, dx), c in zip(taps_y, filt[6:]):
            s = s + (LV[2 + dy:2 + dy + h, 2 + dx:2 + dx + w] - lc) * int(c)
        orig = st.P[:, 8:8 + w].astype(np.int64).copy()
        want = _apply_ll(np.clip((s + 64) >> 7, 0, (1 << bd) - 1), orig, ll, w, h)
        p, lpf, lpfb = st.ptrs()
        lib.oracle_lr_ns_wiener_uv(p, st.stride, _p(st.left), lpf, lpfb, w, h, _p(filt),
                                   lu.L.ctypes.data + 16, lu.stride, lu.top.ctypes.data + 16,
                                   lu.bot.ctypes.data + 16, 1, 1, ds, edges, _p(ll), bd)
        assert np.array_equal(st.P[:, 8:8 + w].astype(np.int64), want), (w, h, edges, ds)


# ---------------------------------------------------------------- GDF
def _gdf(lib, st, p_rows, lpf, lpfb, left, ref, qp):
    w, h = st.w, st.h
    err = np.zeros((h, w), dtype=np.int8)
    P = np.ascontiguousarray(p_rows, dtype=np.uint16)
    L1 = np.ascontiguousarray(lpf, dtype=np.uint16)
    L2 = np.ascontiguousarray(lpfb, dtype=np.uint16)
    LF = np.ascontiguousarray(left, dtype=np.uint16)
    lib.oracle_lr_gdf_prep(_p(err), w, P.ctypes.data + 16, st.stride, _p(LF), L1.ctypes.data + 16,
                           L2.ctypes.data + 16, w, h, ref, qp, st.edges, st.bd)
    return err


@pytest.mark.parametrize("edges", [0, 15, TOP | LEFT])
@pytest.mark.parametrize("ref,qp", [(0, 3), (2, 0), (4, 5)])
def test_gdf_error_12bit_equals_10bit_of_samples_shifted_down(edges, ref, qp):
    """12-bit samples enter GDF's 10-bit domain one by one (>> 2 of each
    sample, gdf_prep looprestoration.rs:1450-1453): any low two bits leave the
    error map unchanged, which a shift of the differences would not."""
    lib = _lib()
    rng = np.random.default_rng(7000 + edges * 31 + ref * 7 + qp)
    for w, h in [(64, 64), (32, 20)]:
        st = Stripe(rng, w, h, 10, edges, hi=1024)
        base = [st.P, st.lpf, st.lpfb, st.left]
        e10 = _gdf(lib, st, *base, ref, qp)
        st.bd = 12
        low = [rng.integers(0, 4, b.shape).astype(np.uint16) for b in base]
        e12 = _gdf(lib, st, *[b.astype(np.uint16) * 4 + l for b, l in zip(base, low)], ref, qp)
        assert np.array_equal(e10, e12), (w, h)
        assert np.any(e10 != 0)


@pytest.mark.parametrize("edges", [0, 15, TOP | LEFT, BOTTOM | RIGHT])
@pytest.mark.parametrize("ref,qp", [(0, 0), (0, 5), (1, 2), (3, 4), (5, 1)])
def test_gdf_error_ignores_a_dc_shift_and_8bit_equals_10bit_times_4(edges, ref, qp):
    lib = _lib()
    rng = np.random.default_rng(5000 + edges * 31 + ref * 7 + qp)
    for w, h in [(64, 64), (32, 20)]:
        st = Stripe(rng, w, h, 8, edges, hi=201)  # 4 x + 200 stays within 10 bits
        base = [st.P, st.lpf, st.lpfb, st.left]
        e8 = _gdf(lib, st, *base, ref, qp)
        st.bd = 10
        e10 = _gdf(lib, st, *[b.astype(np.uint16) * 4 for b in base], ref, qp)
        assert np.array_equal(e8, e10), (w, h, "8-bit vs 10-bit x 4")
        shifted = _gdf(lib, st, *[b.astype(np.uint16) * 4 + 200 for b in base], ref, qp)
        assert np.array_equal(e10, shifted), (w, h, "DC shift")
        assert np.any(e10 != 0)


def _gdf_closed_form(st, ref, qp):
    """gdf_prep (looprestoration.rs:1432-1643) in numpy on the stripe's virtual
    image (rows and columns within 6 of the stripe, the reach of GDF_COORDS).
    Samples enter a 10-bit domain (12-bit: >> 2; 8-bit: differences << 2).
    Each 2x2 block of outputs at (2by, 2bx) takes its class and 4 gradient
    features from the 4x4 window of per-sample gradients |2s - a - c| rows
    2by-1 .. 2by+2, columns 2bx-1 .. 2bx+2, directions (dy, dx) = (1,0),
    (0,1), (1,1), (-1,1): class = (g0 <= g1) | (g2 <= g3) << 1.  Three sums
    = bias + sum min(g >> (4 - up), alpha) w over the features (table slots
    18..21) + sum clip(clip(a - m) + clip(b - m), -512, 511) w over the 18
    symmetric taps (a = s[y - dy][x - dx], b = s[y + dy][x + dx], each
    difference clipped to +-alpha); each sum times scale (intra 8, inter 5)
    is rounded (|v| + 2^14) >> 15 with its sign, clipped to [-scale, scale)
    and the three digits form the index into the error table."""
    T = tables_h.tables()
    alpha_t, weight_t = T["R2T_GDF_ALPHA"], T["R2T_GDF_WEIGHT"]
    bias_t = T["R2T_GDF_BIAS"].reshape(36, 3)
    w, h = st.w, st.h
    down = 2 if st.bd == 12 else 0
    up = 2 if st.bd == 8 else 0
    S = st.image(6, 6).astype(np.int64) >> down

    def at(r0, c0, nr, nc):
        return S[6 + r0:6 + r0 + nr, 6 + c0:6 + c0 + nc]

    hb, wb = (h + 1) // 2, w // 2
    nr, nc = 2 * hb + 2, 2 * wb + 2
    feats = []
    for dy, dx in [(1, 0), (0, 1), (1, 1), (-1, 1)]:
        g = np.abs(2 * at(-1, -1, nr, nc) - at(-1 - dy, -1 - dx, nr, nc) - at(-1 + dy, -1 + dx, nr, nc))
        feats.append(sum(g[i:i + 2 * hb:2, j:j + 2 * wb:2] for i in range(4) for j in range(4)))
    cls = (feats[0] <= feats[1]).astype(np.int64) | ((feats[2] <= feats[3]).astype(np.int64) << 1)
    a_base, w_base = ref * 528 + qp * 88, ref * 1584 + qp * 264

    def per_sample(a):
        return np.repeat(np.repeat(a, 2, 0), 2, 1)[:h, :w]

    sums = []
    for i in range(3):
        v = np.full((hb, wb), bias_t[ref * 6 + qp][i], dtype=np.int64)
        for d in range(4):
            k = 18 + d
            v = v + np.minimum(feats[d] >> (4 - up), alpha_t[a_base + k * 4 + cls]) * \
                weight_t[w_base + i * 88 + k * 4 + cls]
        sums.append(per_sample(v))
    c = per_sample(cls)
    m = at(0, 0, h, w)
    for k, (dy, dx) in enumerate(TAPS["GDF_COORDS"]):
        alpha = alpha_t[a_base + k * 4 + c]
        above = np.clip((at(-dy, -dx, h, w) - m) << up, -alpha, alpha)
        below = np.clip((at(dy, dx, h, w) - m) << up, -alpha, alpha)
        t = np.clip(above + below, -512, 511)
        for i in range(3):
            sums[i] = sums[i] + t * weight_t[w_base + i * 88 + k * 4 + c]
    scale = 8 if ref == 0 else 5
    index = np.zeros((h, w), dtype=np.int64)
    for i in range(3):
        sv = sums[i] * scale
        q = np.sign(sv) * ((np.abs(sv) + (1 << 14)) >> 15)
        index = index * (2 * scale) + np.clip(q, -scale, scale - 1) + scale
    if ref == 0:
        return T["R2T_GDF_INTRA_ERROR"][qp * 4096 + index]
    return T["R2T_GDF_INTER_ERROR"][(ref - 1) * 6000 + qp * 1000 + index]


def _diagonal_texture(rng, st):
    """Samples constant along each anti-diagonal (s = f(r + c)), in every
    buffer: the vertical and horizontal gradients are then equal, so the class
    comparisons tie."""
    f = rng.integers(0, 1 << st.bd, size=st.P.shape[0] + st.P.shape[1] + 8)
    for buf, r0 in ((st.P, 0), (st.lpf, 40), (st.lpfb, 80), (st.left, 7)):
        rr, cc = np.indices(buf.shape)
        buf[:] = f[(rr + cc + r0) % len(f)]


@pytest.mark.parametrize("bd", [8, 10, 12])
@pytest.mark.parametrize("edges", [0, 15, TOP | LEFT, BOTTOM | RIGHT])
def test_gdf_error_is_the_closed_form(bd, edges):
    lib = _lib()
    rng = np.random.default_rng(9000 + bd * 17 + edges)
    for ref in range(6):
        for qp in range(6):
            for w, h, kind in [(64, 64, "random"), (32, 20, "low"), (16, 7, "diagonal")]:
                st = Stripe(rng, w, h, bd, edges, hi=(4 << (bd - 8)) if kind == "low" else None)
                if kind == "diagonal":
                    _diagonal_texture(rng, st)
                got = _gdf(lib, st, st.P, st.lpf, st.lpfb, st.left, ref, qp)
                assert np.array_equal(got, _gdf_closed_form(st, ref, qp)), (ref, qp, w, h, kind)


# ---------------------------------------------------------------- CDEF
AV1_CDEF_DIRECTIONS = [[(-1, 1), (-2, 2)], [(0, 1), (-1, 2)], [(0, 1), (0, 2)], [(0, 1), (1, 2)],
                       [(1, 1), (2, 2)], [(1, 0), (2, 1)], [(1, 0), (2, 0)], [(1, 0), (2, -1)]]


def _av1_constrain(diff, threshold, damping):
    if threshold == 0:
        return 0
    adj = max(0, damping - (int(threshold).bit_length() - 1))
    val = min(abs(diff), max(0, threshold - (abs(diff) >> adj)))
    return val if diff >= 0 else -val


def _av1_cdef(img, y0, x0, w, h, pri, sec, dirn, damping, coeff_shift):
    """The AV1 specification's CDEF filter process on the w x h block at
    (y0, x0) of img, every neighbour available."""
    pri_taps = [[4, 2], [3, 3]][(pri >> coeff_shift) & 1]
    sec_taps = [2, 1]
    out = np.zeros((h, w), dtype=np.int64)
    for i in range(h):
        for j in range(w):
            x = int(img[y0 + i, x0 + j])
            total, mx, mn = 0, x, x
            for k in range(2):
                for sign in (-1, 1):
                    dy, dx = AV1_CDEF_DIRECTIONS[dirn][k]
                    p = int(img[y0 + i + sign * dy, x0 + j + sign * dx])
                    total += pri_taps[k] * _av1_constrain(p - x, pri, damping)
                    mx, mn = max(p, mx), min(p, mn)
                    for off in (-2, 2):
                        dy, dx = AV1_CDEF_DIRECTIONS[(dirn + off) & 7][k]
                        s = int(img[y0 + i + sign * dy, x0 + j + sign * dx])
                        total += sec_taps[k] * _av1_constrain(s - x, sec, damping)
                        mx, mn = max(s, mx), min(s, mn)
            out[i, j] = min(max(x + ((8 + total - (total < 0)) >> 4), mn), mx)
    return out


@pytest.mark.parametrize("bd", [8, 10, 12])
@pytest.mark.parametrize("size", [8, 4])
def test_cdef_block_filter_is_the_av1_filter_process(bd, size):
    lib = _lib()
    rng = np.random.default_rng(6000 + bd * 3 + size)
    cs = bd - 8
    for it in range(48):
        dirn = it % 8
        # AV1 strength ranges: primary 0..15, secondary {0, 1, 2, 4}, scaled to the bit depth
        pri = int(rng.integers(0, 16)) << cs
        sec = [0, 1, 2, 4][int(rng.integers(0, 4))] << cs
        if pri == 0 and sec == 0:
            pri = 5 << cs
        damping = int(rng.integers(3, 7)) + cs - (1 if size == 4 else 0)
        # smooth texture plus noise so that constrain's damping matters
        yy, xx = np.mgrid[0:size + 4, 0:size + 4]
        img = (((yy * 7 + xx * 5) << cs) + rng.integers(-(12 << cs), 12 << cs, size=yy.shape)
               + (128 << cs)).clip(0, (1 << bd) - 1).astype(np.uint16)
        want = _av1_cdef(img, 2, 2, size, size, pri, sec, dirn, damping, cs)
        # the hook's layout: dst rows of stride size + 4 holding the block and
        # its two right neighbours; left[y] the two samples left of row y; top
        # and bottom two rows each at dst's stride, starting at column -2
        stride = size + 4
        dst = np.zeros((size, stride), dtype=np.uint16)
        dst[:, :size + 2] = img[2:2 + size, 2:4 + size]
        left = np.ascontiguousarray(img[2:2 + size, 0:2])
        top = np.zeros((2, stride), dtype=np.uint16)
        bottom = np.zeros((2, stride), dtype=np.uint16)
        top[:, :size + 4] = img[0:2, 0:size + 4]
        bottom[:, :size + 4] = img[2 + size:4 + size, 0:size + 4]
        lib.oracle_cdef_filter_block_16(_p(dst), stride, _p(left), _p(top), _p(bottom), pri, sec, dirn,
                                        damping, size, size, 15, cs)
        dst = dst[:, :size]
        assert np.array_equal(dst.astype(np.int64), want), (it, pri, sec, dirn, damping)


# ---------------------------------------------------------------- deblock
def _deblock(lines, q_thr, side_thr, mwp, mwn, pos_ll=False, neg_ll=False, bd=10):
    """deblock_bd on a 4 x 16 segment (edge between columns 7 and 8)."""
    lib = _lib()
    buf = np.ascontiguousarray(lines, dtype=np.uint16).copy()
    lib.oracle_deblock_edge_16(_p(buf), 8, (1 << bd) - 1, q_thr, side_thr, 16, 1, mwp, mwn, int(pos_ll),
                               int(neg_ll))
    return buf.astype(np.int64)


WIDTHS = [1, 3, 4, 6, 8]


def test_deblock_mirror_symmetry():
    # reversing the samples across the edge (equal widths on both sides)
    # mirrors the output, up to the round-half-up of the corrections
    # ((x + 1024) >> 11 added on the negative side, subtracted on the positive
    # one): on an exact tie the mirrored sample is one higher on the negative
    # side and one lower on the positive side, never anything else
    rng = np.random.default_rng(7000)
    ties = total = 0
    for it in range(600):
        mw = WIDTHS[it % len(WIDTHS)]
        base = rng.integers(300, 700)
        step = rng.integers(-40, 40)
        lines = base + np.where(np.arange(16) >= 8, step, 0) + rng.integers(-6, 7, size=(4, 16))
        q, s = int(rng.integers(1, 40)), int(rng.integers(1, 200))
        a = _deblock(lines, q, s, mw, mw)
        b = _deblock(lines[:, ::-1], q, s, mw, mw)[:, ::-1]
        d = a - b
        assert set(np.unique(d[:, :8])) <= {0, 1} and set(np.unique(d[:, 8:])) <= {0, -1}, (it, mw, q, s, d)
        ties += int(np.count_nonzero(d))
        total += d.size
    assert ties < total // 50


def test_deblock_dc_shift_and_linear_ramps():
    rng = np.random.default_rng(7100)
    for it in range(400):
        mw = WIDTHS[it % len(WIDTHS)]
        mwn = min(mw, [1, 3, 4, 6, 8][int(rng.integers(0, 5))])
        lines = 400 + np.where(np.arange(16) >= 8, rng.integers(-30, 30), 0) + rng.integers(-5, 6, size=(4, 16))
        q, s = int(rng.integers(1, 40)), int(rng.integers(1, 200))
        c = int(rng.integers(-200, 200))
        assert np.array_equal(_deblock(lines + c, q, s, mw, mwn), _deblock(lines, q, s, mw, mwn) + c)
        # a linear ramp along the filter direction is left as it is
        ramp = 300 + int(rng.integers(-9, 10)) * np.arange(16)[None, :] + np.zeros((4, 1), dtype=np.int64)
        assert np.array_equal(_deblock(ramp, q, s, mw, mwn), ramp)


def test_deblock_corrections_taper_away_from_the_edge():
    # on a step the correction is largest next to the edge and never grows
    # outwards; samples beyond the filter width are untouched; the jump shrinks
    for mw in WIDTHS:
        for h in range(2, 30, 3):
            lines = np.tile(np.where(np.arange(16) >= 8, 500 + h, 500), (4, 1))
            out = _deblock(lines, 40, 400, mw, mw)
            corr = np.abs(out[0] - lines[0])
            assert np.all(np.diff(corr[:8]) >= 0) and np.all(np.diff(corr[8:]) <= 0), (mw, h, out[0])
            assert np.array_equal(out[0][:8 - mw], lines[0][:8 - mw]), (mw, h)
            assert np.array_equal(out[0][8 + mw:], lines[0][8 + mw:]), (mw, h)
            assert abs(out[0][8] - out[0][7]) < h, (mw, h)


def test_deblock_reference_known_answers():
    # deblock.rs:2409-2427: a uniform segment is unchanged
    flat = np.full((4, 16), 128)
    assert np.array_equal(_deblock(flat, 10, 20, 3, 3, bd=8), flat)
    # deblock.rs:2429-2461: a sharp 50 -> 200 edge with q = side = 200, width 1, is modified
    step = np.tile(np.where(np.arange(16) >= 8, 200, 50), (4, 1))
    out = _deblock(step, 200, 200, 1, 1, bd=8)
    assert out[0][8] != 200
    # deblock.rs:2463-2490: both sides lossless leave the samples alone
    assert np.array_equal(_deblock(step, 200, 200, 1, 1, True, True, bd=8), step)
    # one lossless side: that side is untouched, the other still filtered
    pos = _deblock(step, 200, 200, 1, 1, True, False, bd=8)
    assert np.array_equal(pos[:, 8:], step[:, 8:]) and not np.array_equal(pos[:, :8], step[:, :8])
    neg = _deblock(step, 200, 200, 1, 1, False, True, bd=8)
    assert np.array_equal(neg[:, :8], step[:, :8]) and not np.array_equal(neg[:, 8:], step[:, 8:])
    # deblock.rs:2392-2406: a sharp 50 -> 200 edge gets a filter width of at
    # most 1 (q 10, side 20): only the two samples next to the edge change
    out = _deblock(step, 10, 20, 3, 3, bd=8)
    assert np.array_equal(out[:, :7], step[:, :7]) and np.array_equal(out[:, 9:], step[:, 9:])
