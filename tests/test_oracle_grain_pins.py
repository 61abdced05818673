"""Pin of the oracle's film-grain application (fgy_32x32xn / fguv_32x32xn,
filmgrain.rs:248-592, driven per row band by apply_grain_row, :688-800).

The grain field the oracle adds is read back from its output, not from its
code: with a flat input and a flat scaling function K, the display sample is
input + round2(K g, 8), and two runs (K = 128 and K = 192) determine every
grain value g exactly.  That field must equal the field built here from the
grain templates (generate_grain_y/uv, pinned separately against the AV1
specification's grain synthesis) and the reference's description of where a
block takes its template window: per row band a 16-bit LFSR seeded from the
band number, per block a 9-bit draw scaled by 3 - block_size (>> 6) for each
of y and x followed by three 16-bit draws, the window at 3 + (2 >> sub)(3 +
offset), and the overlap blends with weights 27/17 (luma) and 23/22
(subsampled chroma) rounded by round2(., 5) and clipped to the grain range.
A second check feeds a textured picture with real scaling points, restricted
and full-range clipping and chroma-from-luma indexing through the oracle and
compares it with input + round2(scaling[index] g, shift) clipped, the index
being the sample (luma) or the combination of the collocated luma average and
the chroma sample (chroma), with scaling from generate_scaling (pinned by the
reference's known answers)."""
import ctypes as C

import numpy as np
import pytest

import rav2d_amd as R
import oracle_lib as O

GW, GH = 82, 73


def _lfsr(bits, s):
    bit = (s ^ (s >> 1) ^ (s >> 3) ^ (s >> 12)) & 1
    s = (s >> 1) | (bit << 15)
    return (s >> (16 - bits)) & ((1 << bits) - 1), s


def _band_offsets(seed, band, n_blocks, block_size):
    """(y, x) template offsets of every block of a row band."""
    s = seed ^ ((((band * 37 + 178) & 0xFF)) << 8) ^ ((band * 173 + 105) & 0xFF)
    out = []
    for _ in range(n_blocks):
        o = []
        for _n in range(2):
            v, s = _lfsr(9, s)
            o.append(((3 - block_size) * v) >> 6)
            for _k in range(3):
                _, s = _lfsr(16, s)
        out.append(o)
    return out


def _round2(x, sh):
    return (x + ((1 << sh) >> 1)) >> sh


def _templates(fgd, seed, bd):
    L = O.load()
    y = np.zeros((GH, GW), dtype=np.int16)
    L.oracle_generate_grain_y(O.ptr(y), C.byref(fgd), seed, bd)
    uv = []
    for c in range(2):
        t = np.zeros((GH, GW), dtype=np.int16)
        L.oracle_generate_grain_uv(O.ptr(t), O.ptr(y), C.byref(fgd), seed, c, 1, 1, bd)
        uv.append(t)
    return [y.astype(np.int64)] + [t.astype(np.int64) for t in uv]


def _field(lut, pw, luma_h, sub, seed, block_size, overlap, bd):
    """Grain value of every sample of a plane (pw wide; luma_h luma rows)."""
    bs = 16 << block_size
    pbs = bs >> sub
    gmax = (128 << (bd - 8)) - 1
    gmin = -(128 << (bd - 8))
    wh = [[27, 17], [17, 27]] if not sub else [[23, 22], [0, 0]]
    n_bands = (luma_h + bs - 1) // bs
    n_blocks = (pw + pbs - 1) // pbs
    ph = (luma_h + sub) >> sub
    G = np.zeros((ph, pw), dtype=np.int64)
    offs = [_band_offsets(seed, r, n_blocks, block_size) for r in range(n_bands)]

    def win(o, ly, lx, by, bx):
        oy = 3 + (2 >> sub) * (3 + o[0])
        ox = 3 + (2 >> sub) * (3 + o[1])
        return lut[oy + ly + pbs * by, ox + lx + pbs * bx]

    def blend(a, b, wa, wb):
        return min(max(_round2(a * wa + b * wb, 5), gmin), gmax)

    for r in range(n_bands):
        bh = (min(bs, luma_h - r * bs) + sub) >> sub
        y0 = r * pbs
        for b in range(n_blocks):
            x0 = b * pbs
            bw = min(pbs, pw - x0)
            xs = min(2 >> sub, bw) if overlap and b > 0 else 0
            ys = min(2 >> sub, bh) if overlap and r > 0 else 0
            for ly in range(bh):
                for lx in range(bw):
                    g = int(win(offs[r][b], ly, lx, 0, 0))
                    if lx < xs:
                        g = blend(int(win(offs[r][b - 1], ly, lx, 0, 1)), g, wh[lx][0], wh[lx][1])
                    if ly < ys:
                        top = int(win(offs[r - 1][b], ly, lx, 1, 0))
                        if lx < xs:
                            top = blend(int(win(offs[r - 1][b - 1], ly, lx, 1, 1)), top, wh[lx][0], wh[lx][1])
                        g = blend(top, g, wh[ly][0], wh[ly][1])
                    G[y0 + ly, x0 + lx] = g
    return G


def _params(w, h, bd, seed, block_size, overlap):
    sf = R.SynthFrame(w, h, bd, 0, seed=seed, tools=R.TOOL_GRAIN)
    p = R.FrameParams.from_buffer_copy(sf.params)
    sf.close()
    p.grain_present = 1
    p.grain_seed = (seed * 7919 + 11) & 0xffff
    g = p.grain
    g.block_size, g.overlap_flag = block_size, overlap
    g.clip_to_restricted_range, g.mc_identity = 0, 0
    g.chroma_scaling_from_luma = 0
    return p


def _flat_scaling(p, K):
    g = p.grain
    g.scaling_shift = 8
    for pl in range(3):
        g.num_points[pl] = 2
        g.points[pl][0][0], g.points[pl][0][1] = 0, K
        g.points[pl][1][0], g.points[pl][1][1] = 255, K
    for uv in range(2):
        g.uv_mult[uv] = g.uv_luma_mult[uv] = g.uv_offset[uv] = 0


def _oracle_grain(p, planes):
    L = O.load()
    out = [np.empty_like(a) for a in planes]
    bps = planes[0].itemsize
    rc = L.oracle_apply_grain(C.byref(p), planes[0].ctypes.data, planes[1].ctypes.data, planes[2].ctypes.data,
                              out[0].ctypes.data, out[1].ctypes.data, out[2].ctypes.data,
                              planes[0].shape[1] * bps, planes[1].shape[1] * bps)
    assert rc == 0
    return [o.astype(np.int64) for o in out]


CASES = [
    # (w, h, bitdepth, block_size, overlap)
    (264, 136, 10, 0, 1),
    (264, 136, 10, 1, 1),
    (200, 104, 8, 0, 1),
    (200, 104, 8, 1, 0),
    (136, 72, 10, 0, 0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d_%db_bs%d_ov%d" % c)
def test_grain_field_is_the_template_windows(case):
    w, h, bd, bsz, ov = case
    dt = np.uint16 if bd > 8 else np.uint8
    mid = 128 << (bd - 8)
    planes = [np.full((h, w), mid, dtype=dt), np.full((h // 2, w // 2), mid, dtype=dt),
              np.full((h // 2, w // 2), mid, dtype=dt)]
    runs = []
    for K in (128, 192):
        p = _params(w, h, bd, 100 + bsz * 10 + ov, bsz, ov)
        _flat_scaling(p, K)
        runs.append([o - mid for o in _oracle_grain(p, planes)])
    # noise = round2(K g, 8): K = 128 leaves two candidates, K = 192 picks one
    fields = []
    for n1, n2 in zip(*runs):
        g_hi, g_lo = 2 * n1, 2 * n1 - 1
        pick_hi = _round2(192 * g_hi, 8) == n2
        pick_lo = _round2(192 * g_lo, 8) == n2
        assert np.all(pick_hi | pick_lo) and not np.any(pick_hi & pick_lo)
        fields.append(np.where(pick_hi, g_hi, g_lo))
    luts = _templates(p.grain, p.grain_seed, bd)
    want_y = _field(luts[0], w, h, 0, p.grain_seed, bsz, ov, bd)
    assert np.array_equal(fields[0], want_y), "luma grain field"
    for c in range(2):
        want = _field(luts[1 + c], w // 2, h, 1, p.grain_seed, bsz, ov, bd)
        assert np.array_equal(fields[1 + c], want), "chroma %d grain field" % c


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d_%db_bs%d_ov%d" % c)
@pytest.mark.parametrize("mode", ["full_range", "restricted", "restricted_mc_identity", "cfl"])
def test_grain_output_is_input_plus_scaled_field(case, mode):
    w, h, bd, bsz, ov = case
    rng = np.random.default_rng(hash(case) % 1000 + len(mode))
    dt = np.uint16 if bd > 8 else np.uint8
    hi = (1 << bd) - 1
    yy, xx = np.mgrid[0:h, 0:w]
    luma = ((yy * 3 + xx * 2) % (hi + 1) + rng.integers(0, 40 << (bd - 8), size=(h, w))).clip(0, hi).astype(dt)
    chroma = [rng.integers(0, hi + 1, size=(h // 2, w // 2)).astype(dt) for _ in range(2)]
    planes = [luma] + chroma
    p = _params(w, h, bd, 300 + bsz * 10 + ov, bsz, ov)
    g = p.grain
    g.scaling_shift = int(rng.integers(8, 12))
    for pl in range(3):
        n = 8 if pl == 0 else 5
        xs = np.sort(rng.choice(256, size=n, replace=False))
        g.num_points[pl] = n
        for i in range(n):
            g.points[pl][i][0], g.points[pl][i][1] = int(xs[i]), int(rng.integers(0, 256))
    for uv in range(2):
        g.uv_mult[uv] = int(rng.integers(-128, 128))
        g.uv_luma_mult[uv] = int(rng.integers(-128, 128))
        g.uv_offset[uv] = int(rng.integers(-256, 256))
    g.clip_to_restricted_range = 0 if mode == "full_range" else 1
    g.mc_identity = 1 if mode == "restricted_mc_identity" else 0
    g.chroma_scaling_from_luma = 1 if mode == "cfl" else 0
    got = _oracle_grain(p, planes)

    L = O.load()
    scal = []
    for pl in range(3):
        s = np.zeros(1 << bd, dtype=np.uint8)
        pts = (C.c_uint8 * 28).from_buffer_copy(bytes(g.points[pl]))
        L.oracle_generate_scaling(bd, C.cast(pts, C.c_void_p), g.num_points[pl], O.ptr(s))
        scal.append(s.astype(np.int64))
    luts = _templates(g, p.grain_seed, bd)
    bdm8 = bd - 8
    if g.clip_to_restricted_range:
        lo, top_y, top_c = 16 << bdm8, 235 << bdm8, (235 if g.mc_identity else 240) << bdm8
    else:
        lo, top_y, top_c = 0, hi, hi
    Gy = _field(luts[0], w, h, 0, p.grain_seed, bsz, ov, bd)
    src = luma.astype(np.int64)
    want_y = np.clip(src + _round2(scal[0][src] * Gy, g.scaling_shift), lo, top_y)
    assert np.array_equal(got[0], want_y), "luma"
    avg = (src[0::2, 0::2] + src[0::2, 1::2] + 1) >> 1
    for c in range(2):
        Gc = _field(luts[1 + c], w // 2, h, 1, p.grain_seed, bsz, ov, bd)
        sv = chroma[c].astype(np.int64)
        if g.chroma_scaling_from_luma:
            idx, sc = avg, scal[0]
        else:
            comb = avg * g.uv_luma_mult[c] + sv * g.uv_mult[c]
            idx, sc = np.clip((comb >> 6) + g.uv_offset[c] * (1 << bdm8), 0, hi), scal[1 + c]
        want = np.clip(sv + _round2(sc[idx] * Gc, g.scaling_shift), lo, top_c)
        assert np.array_equal(got[1 + c], want), "chroma %d" % c
